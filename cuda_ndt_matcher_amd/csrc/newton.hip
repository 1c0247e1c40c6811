// Persistent Newton kernel and scoring kernels (gfx950).
//
// One launch runs whole alignments (NdtOptimizer::align,
// src/ndt_cuda/src/optimization/solver.rs:135-305) for `slots` independent
// requests.  Each slot owns `wg_per_slot` workgroups.  Per Newton iteration:
//   - every thread transforms its points (f64), finds the neighbouring voxels
//     through the hash, and accumulates score, gradient and the lower
//     triangle of the Hessian in registers (compute_derivatives_cpu_with_metric,
//     derivatives/cpu.rs:417-538);
//   - the workgroup reduces them with wave shuffles and LDS;
//   - with more than one workgroup per slot, the workgroups publish partials
//     and meet at a slot barrier, then every workgroup sums the partials in
//     rank order, so all of them hold bit-identical totals;
//   - thread 0 of every workgroup runs the same FP64 Newton step
//     (newton.rs:160-176), convergence test, step clamp or More-Thuente
//     line search (more_thuente.rs:268-420) and pose update, so no second
//     barrier is needed.
// After the loop the slot computes NVTL at the final pose with the Isometry
// transform (scoring/nvtl.rs:211-268) and rank 0 writes the result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "ndt_device.hpp"
#include "ndt_internal.hpp"

namespace ndt {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxIter = 256;
constexpr int kPartialStride = 32;  // doubles per workgroup partial
constexpr int kRankGroups = kBlock / 16;  // slot_reduce: 16 column pairs x kRankGroups rank groups
constexpr int kReduceLds = kRankGroups * kPartialStride;
static_assert(kReduceLds >= (kBlock / kWave) * kAccSlots, "slot_reduce scratch");
static_assert(kAccSlots <= kPartialStride, "partial stride");

// ---------------------------------------------------------------------------
// More-Thuente line search as a state machine fed one (score, derivative)
// evaluation at a time (optimization/more_thuente.rs:268-420, defaults
// step_max 0.1, step_min 1e-9, mu 1e-4, nu 0.9, 10 iterations).
struct MTState {
  double phi0, dphi0, a_t, phi_t, dphi_t, psi_t, dpsi_t, a_l, f_l, g_l, a_u, f_u, g_u;
  int open, it, phase, done, conv;
};

constexpr double kMtMax = 0.1, kMtMin = 1e-9, kMtMu = 1e-4, kMtNu = 0.9;
constexpr int kMtIters = 10;

__device__ double mt_trial(double a_l, double f_l, double g_l, double a_u, double f_u, double g_u, double a_t,
                           double f_t, double g_t) {
  if (f_t > f_l) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    if (w_sq < 0.0) return a_l - 0.5 * (a_l - a_t) * g_l / (g_l - (f_l - f_t) / (a_l - a_t));
    const double w = sqrt(w_sq);
    const double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    const double denom = g_l - (f_l - f_t) / (a_l - a_t);
    if (fabs(denom) < 1e-15) return a_c;
    const double a_q = a_l - 0.5 * (a_l - a_t) * g_l / denom;
    if (fabs(a_c - a_l) < fabs(a_q - a_l)) return a_c;
    return 0.5 * (a_q + a_c);
  }
  if (g_t * g_l < 0.0) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    if (w_sq < 0.0) return a_l - (a_l - a_t) / (g_l - g_t) * g_l;
    const double w = sqrt(w_sq);
    const double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    const double a_s = a_l - (a_l - a_t) / (g_l - g_t) * g_l;
    return (fabs(a_c - a_t) >= fabs(a_s - a_t)) ? a_c : a_s;
  }
  if (fabs(g_t) <= fabs(g_l)) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    double a_c;
    if (w_sq >= 0.0) {
      const double w = sqrt(w_sq);
      a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    } else {
      a_c = (a_t > a_l) ? a_t + 0.66 * (a_u - a_t) : a_t - 0.66 * (a_t - a_u);
    }
    const double a_s = (fabs(g_l - g_t) > 1e-15) ? a_l - (a_l - a_t) / (g_l - g_t) * g_l : a_c;
    const double nxt = (fabs(a_c - a_t) < fabs(a_s - a_t)) ? a_c : a_s;
    if (a_t > a_l) return fmin(a_t + 0.66 * (a_u - a_t), nxt);
    return fmax(a_t + 0.66 * (a_u - a_t), nxt);
  }
  const double z = 3.0 * (f_t - f_u) / (a_t - a_u) - g_t - g_u;
  const double w_sq = z * z - g_t * g_u;
  if (w_sq < 0.0) return 0.5 * (a_l + a_u);
  const double w = sqrt(w_sq);
  return a_u + (a_t - a_u) * (w - g_u - z) / (g_t - g_u + 2.0 * w);
}

__device__ bool mt_update(MTState& s, double f_t, double g_t) {
  if (f_t > s.f_l) {
    s.a_u = s.a_t; s.f_u = f_t; s.g_u = g_t;
    return false;
  }
  if (g_t * (s.a_l - s.a_t) > 0.0) {
    s.a_l = s.a_t; s.f_l = f_t; s.g_l = g_t;
    return false;
  }
  if (g_t * (s.a_l - s.a_t) < 0.0) {
    s.a_u = s.a_l; s.f_u = s.f_l; s.g_u = s.g_l;
    s.a_l = s.a_t; s.f_l = f_t; s.g_l = g_t;
    return false;
  }
  return true;
}

__device__ inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ void mt_begin(MTState& s, double score0, double deriv0, double init_step) {
  s.phi0 = -score0;
  s.dphi0 = -deriv0;
  s.a_t = clampd(init_step, kMtMin, kMtMax);
  s.phase = 0;
  s.done = 0;
  s.conv = 0;
}

// Consumes the evaluation at s.a_t; either finishes or sets the next a_t.
__device__ void mt_feed(MTState& s, double score_t, double deriv_t) {
  s.phi_t = -score_t;
  s.dphi_t = -deriv_t;
  if (s.phase == 0) {
    s.a_l = 0.0;
    s.a_u = 0.0;
    s.f_l = s.phi0 - s.phi0 - kMtMu * s.dphi0 * 0.0;
    s.g_l = s.dphi0 - kMtMu * s.dphi0;
    s.f_u = s.f_l;
    s.g_u = s.g_l;
    s.open = 1;
    s.it = 0;
    s.psi_t = s.phi_t - s.phi0 - kMtMu * s.dphi0 * s.a_t;
    s.dpsi_t = s.dphi_t - kMtMu * s.dphi0;
  } else {
    s.psi_t = s.phi_t - s.phi0 - kMtMu * s.dphi0 * s.a_t;
    s.dpsi_t = s.dphi_t - kMtMu * s.dphi0;
    if (s.open && s.psi_t <= 0.0 && s.dpsi_t >= 0.0) {
      s.open = 0;
      s.f_l = s.f_l + s.phi0 - kMtMu * s.dphi0 * s.a_l;
      s.g_l += kMtMu * s.dphi0;
      s.f_u = s.f_u + s.phi0 - kMtMu * s.dphi0 * s.a_u;
      s.g_u += kMtMu * s.dphi0;
    }
    const bool stop = s.open ? mt_update(s, s.psi_t, s.dpsi_t) : mt_update(s, s.phi_t, s.dphi_t);
    if (stop) {
      s.done = 1;
      s.conv = 0;
      return;
    }
    s.it += 1;
  }
  if (s.it >= kMtIters) {
    s.done = 1;
    s.conv = 0;
    return;
  }
  if (s.psi_t <= 0.0 && fabs(s.dphi_t) <= kMtNu * fabs(s.dphi0)) {
    s.done = 1;
    s.conv = 1;
    return;
  }
  const double nt = s.open ? mt_trial(s.a_l, s.f_l, s.g_l, s.a_u, s.f_u, s.g_u, s.a_t, s.psi_t, s.dpsi_t)
                           : mt_trial(s.a_l, s.f_l, s.g_l, s.a_u, s.f_u, s.g_u, s.a_t, s.phi_t, s.dphi_t);
  s.a_t = clampd(nt, kMtMin, kMtMax);
  s.phase = 1;
}

// ---------------------------------------------------------------------------
// Per-point work.  R is the arithmetic type of the per-point terms: double
// for the FP64 form, float for the FP32 form.
template <class R>
__device__ inline void point_jacobian(const PoseTables& T, R px, R py, R pz, R Jr[3][3], R ht[15], bool hess) {
  R gt[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) gt[r] = (R)T.j[r][0] * px + (R)T.j[r][1] * py + (R)T.j[r][2] * pz;
  Jr[0][0] = 0;     Jr[1][0] = gt[0]; Jr[2][0] = gt[1];
  Jr[0][1] = gt[2]; Jr[1][1] = gt[3]; Jr[2][1] = gt[4];
  Jr[0][2] = gt[5]; Jr[1][2] = gt[6]; Jr[2][2] = gt[7];
  if (hess) {
#pragma unroll
    for (int r = 0; r < 15; ++r) ht[r] = (R)T.h[r][0] * px + (R)T.h[r][1] * py + (R)T.h[r][2] * pz;
  } else {
#pragma unroll
    for (int r = 0; r < 15; ++r) ht[r] = 0;
  }
}

// Per-thread accumulation of the 28 sums.  FP64: straight into acc.  FP32:
// compensated (two-sum) float pairs, added to acc in double at the end of
// the pass; a thread can hold many points (batched launches), and the
// compensation keeps their float sum exact to about 2^-48 relative.
template <class R>
struct PointSums;

template <>
struct PointSums<double> {
  double* acc;
  __device__ explicit PointSums(double* a) : acc(a) {}
  __device__ void add(int i, double x) { acc[i] += x; }
  __device__ void finish() {}
};

template <>
struct PointSums<float> {
  double* acc;
  float hi[kAccVals], lo[kAccVals];
  __device__ explicit PointSums(double* a) : acc(a) {
#pragma unroll
    for (int i = 0; i < kAccVals; ++i) hi[i] = lo[i] = 0.0f;
  }
  __device__ void add(int i, float x) {
    const float s = hi[i] + x;
    const float bb = s - hi[i];
    lo[i] += (hi[i] - (s - bb)) + (x - bb);
    hi[i] = s;
  }
  __device__ void finish() {
#pragma unroll
    for (int i = 0; i < kAccVals; ++i) acc[i] += (double)hi[i] + (double)lo[i];
  }
};

__device__ inline double exp_r(double x) { return exp(x); }
__device__ inline float exp_r(float x) { return expf(x); }

// Accumulates the derivative contributions of points first, first + stride,
// ... below n.
// acc: 28 doubles (score, grad, H lower), corr: pair count.
//
// Every pair term of compute_derivative_single (derivatives/cpu.rs:139-229)
// has the form J_i^T X J_j or y . d2T/dp_i dp_j with J the point's 3x6
// Jacobian, so the pairs of one point are first summed into
//   s = sum -d1 e_raw,  a = sum e xc,  M = sum e (C^T - d2 xc xc^T)
// (e = d1 d2 e_raw, xc = x^T C) and the Jacobian is applied once per point:
//   g_j = a . J_j,  H_ij = J_i^T M J_j + a . d2T/dp_i dp_j   (i >= j).
// The point-to-plane metric (cpu.rs:246-329) has the same shape with
// a = sum e d n and M = sum e (1 - d2 d^2) n n^T.
//
// The neighbours of a point are first staged (record and slot) in this
// thread's column of an LDS table and then visited in one loop, so a wave
// runs the pair computation max-over-lanes(neighbours) times rather than
// once per candidate position that any lane accepted.
constexpr int kStageK = 8;  // staged neighbours per thread before a flush
struct Stage {
  float4* rec;   // [kStageK][3][kBlock]
  int32_t* slot; // [kStageK][kBlock]
};

template <bool kHess, class R = double>
__device__ inline void derivative_points(const GridView& g, const NewtonConfig& cfg, const PoseTables& T,
                                         const float* __restrict__ pts, uint32_t n, uint32_t first,
                                         uint32_t stride, double* acc, uint32_t& corr, const Stage& st) {
  const R d1 = (R)cfg.d1, d2 = (R)cfg.d2;
  const uint32_t tid = threadIdx.x;
  PointSums<R> sums(acc);
  for (uint32_t i = first; i < n; i += stride) {
    // Keeps the pose tables in LDS until they are used: a thread usually has
    // one point, and hoisting the 70 table entries out of this loop would
    // hold them in registers through the whole pass.
    asm volatile("" ::: "memory");
    const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    double t[3];
    t[0] = T.R[0] * px + T.R[1] * py + T.R[2] * pz + T.t[0];
    t[1] = T.R[3] * px + T.R[4] * py + T.R[5] * pz + T.t[1];
    t[2] = T.R[6] * px + T.R[7] * py + T.R[8] * pz + T.t[2];
    R sc = 0, a0 = 0, a1 = 0, a2 = 0;
    R M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = 0;
    uint32_t np = 0;
    // One pair's terms: sraw = e_raw, and the a and M increments in the
    // common form a += w u, M += w C^T (point-to-distribution only) +
    // q u u^T.  ok is false when the pair is counted but contributes zero.
    auto eval = [&](int32_t s, const VoxelRecord& rec, R& sraw, R& w, R& q, R u[3], bool& ok) {
      // residual in double (the point is placed in double), then R
      const R x0 = (R)(t[0] - (double)rec.m0.x), x1 = (R)(t[1] - (double)rec.m0.y), x2 = (R)(t[2] - (double)rec.m0.z);
      if (cfg.metric == 1) {
        u[0] = g.axis[3 * s];
        u[1] = g.axis[3 * s + 1];
        u[2] = g.axis[3 * s + 2];
        const R d = x0 * u[0] + x1 * u[1] + x2 * u[2];
        const R dsq = d * d;
        const R e_raw = exp_r(-d2 * dsq * (R)0.5);
        const R nz = d2 * e_raw;
        ok = (nz >= (R)0 && nz <= (R)1) && isfinite(nz);
        const R e = d1 * d2 * e_raw;
        sraw = e_raw;
        w = e * d;
        q = e * (-d2 * dsq + (R)1);
      } else {
        const float4 m0 = rec.m0, m1 = rec.m1, m2 = rec.m2;
        const R C[9] = {m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
        u[0] = x0 * C[0] + x1 * C[3] + x2 * C[6];
        u[1] = x0 * C[1] + x1 * C[4] + x2 * C[7];
        u[2] = x0 * C[2] + x1 * C[5] + x2 * C[8];
        const R xcx = u[0] * x0 + u[1] * x1 + u[2] * x2;
        const R e_raw = exp_r(-d2 * xcx * (R)0.5);
        const R nz = d2 * e_raw;
        ok = (nz >= (R)0 && nz <= (R)1) && isfinite(nz);
        const R e = nz * d1;
        sraw = e_raw;
        w = e;
        q = -d2 * e;
      }
    };
    auto accum = [&](const VoxelRecord& rec, R sraw, R w, R q, const R u[3]) {
      sc += -d1 * sraw;
      a0 += w * u[0];
      a1 += w * u[1];
      a2 += w * u[2];
      if (kHess) {
        if (cfg.metric == 1) {
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[3 * r + c] += q * (u[r] * u[c]);
        } else {
          const float4 m0 = rec.m0, m1 = rec.m1, m2 = rec.m2;
          const R C[9] = {m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[3 * r + c] += w * C[3 * c + r] + q * (u[r] * u[c]);
        }
      }
    };
    auto staged_rec = [&](uint32_t v) {
      return VoxelRecord{st.rec[(3 * v) * kBlock + tid], st.rec[(3 * v + 1) * kBlock + tid],
                         st.rec[(3 * v + 2) * kBlock + tid]};
    };
    // Neighbours kStageK at a time, in enumeration order; a point with more
    // (rare: at most 27 exist) enumerates again for the next ones.
    for (uint32_t skip = 0;; skip += kStageK) {
      uint32_t seen = 0, staged = 0;
      for_each_neighbor(g, cfg.search_mode, (float)t[0], (float)t[1], (float)t[2],
                        [&](int32_t s, VoxelRecord rec) {
        if (seen >= skip && staged < kStageK) {
          st.rec[(3 * staged) * kBlock + tid] = rec.m0;
          st.rec[(3 * staged + 1) * kBlock + tid] = rec.m1;
          st.rec[(3 * staged + 2) * kBlock + tid] = rec.m2;
          st.slot[staged * kBlock + tid] = s;
          ++staged;
        }
        ++seen;
      });
      corr += staged;
      np += staged;
      // Two pairs per step: their terms are independent dependency chains
      // (the exponential above all), so evaluating both before accumulating
      // either overlaps them; the sums still take the pairs in order.
      uint32_t v = 0;
      for (; v + 1 < staged; v += 2) {
        const VoxelRecord r0 = staged_rec(v), r1 = staged_rec(v + 1);
        R s0, w0, q0, u0[3], s1, w1, q1, u1[3];
        bool ok0, ok1;
        eval(st.slot[v * kBlock + tid], r0, s0, w0, q0, u0, ok0);
        eval(st.slot[(v + 1) * kBlock + tid], r1, s1, w1, q1, u1, ok1);
        if (ok0) accum(r0, s0, w0, q0, u0);
        if (ok1) accum(r1, s1, w1, q1, u1);
      }
      if (v < staged) {
        const VoxelRecord r0 = staged_rec(v);
        R s0, w0, q0, u0[3];
        bool ok0;
        eval(st.slot[v * kBlock + tid], r0, s0, w0, q0, u0, ok0);
        if (ok0) accum(r0, s0, w0, q0, u0);
      }
      if (seen <= skip + kStageK) break;
    }
    if (np == 0) continue;
    R Jr[3][3], ht[15];
    point_jacobian<R>(T, (R)px, (R)py, (R)pz, Jr, ht, kHess);
    sums.add(0, sc);
    sums.add(1, a0);
    sums.add(2, a1);
    sums.add(3, a2);
#pragma unroll
    for (int c = 0; c < 3; ++c) sums.add(4 + c, a0 * Jr[0][c] + a1 * Jr[1][c] + a2 * Jr[2][c]);
    if (kHess) {
      R MJ[3][3];  // M * Jr
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) MJ[r][c] = M[3 * r] * Jr[0][c] + M[3 * r + 1] * Jr[1][c] + M[3 * r + 2] * Jr[2][c];
      // translation block
      sums.add(7 + tri(0, 0), M[0]);
      sums.add(7 + tri(1, 0), M[3]);
      sums.add(7 + tri(1, 1), M[4]);
      sums.add(7 + tri(2, 0), M[6]);
      sums.add(7 + tri(2, 1), M[7]);
      sums.add(7 + tri(2, 2), M[8]);
      // rotation rows against translation columns: sum_k Jr[k][i] M[k][j]
#pragma unroll
      for (int ri = 0; ri < 3; ++ri)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          sums.add(7 + tri(3 + ri, j), Jr[0][ri] * M[j] + Jr[1][ri] * M[3 + j] + Jr[2][ri] * M[6 + j]);
      // rotation block plus the point-Hessian terms a . d2T/dp_i dp_j
      const R pa = a1 * ht[0] + a2 * ht[1];
      const R pb = a1 * ht[2] + a2 * ht[3];
      const R pc = a1 * ht[4] + a2 * ht[5];
      const R pd = a0 * ht[6] + a1 * ht[7] + a2 * ht[8];
      const R pe = a0 * ht[9] + a1 * ht[10] + a2 * ht[11];
      const R pf = a0 * ht[12] + a1 * ht[13] + a2 * ht[14];
      const R ph[3][3] = {{pa, pb, pc}, {pb, pd, pe}, {pc, pe, pf}};
#pragma unroll
      for (int ri = 0; ri < 3; ++ri)
#pragma unroll
        for (int rj = 0; rj <= ri; ++rj)
          sums.add(7 + tri(3 + ri, 3 + rj),
                   Jr[0][ri] * MJ[0][rj] + Jr[1][ri] * MJ[1][rj] + Jr[2][ri] * MJ[2][rj] + ph[ri][rj]);
    }
  }
  sums.finish();
}

__device__ inline double point_score(const double x[3], const float4& m0, const float4& m1, const float4& m2,
                                     double d1, double d2) {
  // compute_point_voxel_score (scoring/nvtl.rs:196-206)
  const double d[3] = {x[0] - (double)m0.x, x[1] - (double)m0.y, x[2] - (double)m0.z};
  const double C[9] = {m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
  double Cd[3];
  for (int r = 0; r < 3; ++r) Cd[r] = C[3 * r] * d[0] + C[3 * r + 1] * d[1] + C[3 * r + 2] * d[2];
  const double m = d[0] * Cd[0] + d[1] * Cd[1] + d[2] * Cd[2];
  return -d1 * exp(-d2 * m / 2.0);
}

// NVTL contributions (max per point, points with neighbours) over a strided range.
__device__ inline void nvtl_points(const GridView& g, double d1, double d2, const double q[4], const double tr[3],
                                   const float* __restrict__ pts, uint32_t n, uint32_t first, uint32_t stride,
                                   double& sum, double& cnt) {
  for (uint32_t i = first; i < n; i += stride) {
    double x[3];
    iso_transform(q, tr, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], x);
    double best = -HUGE_VAL;
    bool any = false;
    for_each_radius(g, (float)x[0], (float)x[1], (float)x[2], [&](int32_t, VoxelRecord r) {
      const double sc = point_score(x, r.m0, r.m1, r.m2, d1, d2);
      best = fmax(best, sc);
      any = true;
    });
    if (any) {
      sum += best;
      cnt += 1.0;
    }
  }
}

// ---------------------------------------------------------------------------
struct SlotState {
  double pose[6];
  double grad[6];
  double delta[6];
  double score;
  double best_pose[6];
  double dir[6];
  double last_H[36];
  double best_score;
  double step;
  double final_pose[6];
  double final_score;
  double final_H[36];
  double hist[3 * (kMaxIter + 1)];
  MTState mt;
  int nhist;
  int status;       // -1 running
  int iterations;
  int ctrl;         // 0: take a step, 1: stop, 2: line search, 3: break to the max-iterations result
  int need_solve;
  uint32_t last_corr, final_corr;
  uint32_t corr_total;  // pairs summed over the derivative passes with the Hessian
};

// Slot-wide exchange state, identical in every workgroup of a slot.
//
// Two protocols.  Write-once (exchanges > 0, the one-shot launches): exchange
// e of a launch uses its own buffer, filled with kExchangeEmpty before the
// launch; each workgroup stores its partials there with relaxed agent-scope
// atomic stores and every workgroup reads all of them with relaxed atomic
// loads, reloading a value until it is no longer the sentinel.  A location
// is written once per launch, so a value that is not the sentinel is the
// exchange's own and no fence or arrival counter is needed.
//
// Alternating buffers with arrival counters (exchanges == 0: the server,
// whose slots run any number of jobs).  Arrivals are counted in two levels: each lane that stores a partial value
// adds to the counter of its workgroup's group of kArrivalGroup ranks (its
// own release covers its own store), and the lane that completes a group
// adds 1 to the slot's top counter, which every workgroup polls.  (One
// arrival per workgroup instead of one per value measured the same.)
// Counters sit in separate 128-byte lines.
constexpr int kArrivalGroup = 16;
constexpr int kCounterStride = 32;  // uint32 per counter line
// Bits of the sentinel that marks an exchange value not yet written: a
// signalling NaN, which no arithmetic produces.
constexpr unsigned long long kExchangeEmpty = 0x7FF0DEADBEEF0001ull;

struct SlotSync {
  double* partials;         // E x W x kPartialStride doubles (write-once) or 2 x W x kPartialStride (alternating)
  int E;                    // write-once buffers per slot (0: alternating protocol)
  uint32_t* top;            // slot counter: completed groups
  uint32_t* group;          // this rank's group counter: arrivals
  uint32_t group_size;      // workgroups in this rank's group
  uint32_t groups;          // groups in the slot
  uint32_t gen;             // reductions done
  uint32_t group_arrivals;  // group counter value that completes the current reduction
  uint32_t top_arrivals;    // top counter value that completes the current reduction
  int W;
  int rank;
};

__host__ __device__ constexpr int counter_lines(int W) { return 1 + (W + kArrivalGroup - 1) / kArrivalGroup; }

template <int N>
__device__ inline void slot_reduce(double (&v)[N], int count, double* lds, double* red, double* tot, SlotSync& sy,
                                   unsigned long long* stamp = nullptr) {
  const int t = threadIdx.x;
  double wg;
  if constexpr (N == kAccSlots) {
    block_sums_lds<kBlock>(v, count, red, lds);
    wg = t < count ? lds[t] : 0.0;
    __syncthreads();
  } else {
    wg = block_sums<kBlock, N>(v, count, lds);
  }
  if (stamp && t == 0) *stamp = wall_clock64();
  const int W = sy.W;
  if (W == 1) {
    if (t < count) tot[t] = wg;
    __syncthreads();
    return;
  }
  if (sy.E > 0) {
    // Write-once buffer of this exchange (the host sizes E for every
    // exchange of an alignment; the modulo only keeps the index in bounds).
    unsigned long long* wb =
        reinterpret_cast<unsigned long long*>(sy.partials + (size_t)(sy.gen % (uint32_t)sy.E) * W * kPartialStride);
    ++sy.gen;
    if (t < count)
      __hip_atomic_store(wb + (size_t)sy.rank * kPartialStride + t, (unsigned long long)__double_as_longlong(wg),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Same reading order and summation as the alternating protocol below.
    const int pair = t & 15, grp = t >> 4;
    double s0 = 0.0, s1 = 0.0;
    if (2 * pair < count) {
      const bool second = 2 * pair + 1 < count;
      const unsigned long long* col = wb + 2 * pair;
      for (int r0 = grp; r0 < W; r0 += 8 * kRankGroups) {
        unsigned long long x[8][2];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int r = r0 + k * kRankGroups;
          const unsigned long long* c = col + (size_t)r * kPartialStride;
          x[k][0] = r < W ? __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
          x[k][1] = (r < W && second) ? __hip_atomic_load(c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int r = r0 + k * kRankGroups;
          const unsigned long long* c = col + (size_t)r * kPartialStride;
          // Every workgroup of the slot is resident (cooperative launch) and
          // stores its values in this exchange, so each wait ends.
          while (x[k][0] == kExchangeEmpty) {
            __builtin_amdgcn_s_sleep(1);
            x[k][0] = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          while (x[k][1] == kExchangeEmpty) {
            __builtin_amdgcn_s_sleep(1);
            x[k][1] = __hip_atomic_load(c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          s0 += __longlong_as_double((long long)x[k][0]);
          s1 += __longlong_as_double((long long)x[k][1]);
        }
      }
    }
    lds[grp * kPartialStride + 2 * pair] = s0;
    lds[grp * kPartialStride + 2 * pair + 1] = s1;
    __syncthreads();
    if (t < count) {
      double s = 0.0;
#pragma unroll
      for (int g = 0; g < kRankGroups; ++g) s += lds[g * kPartialStride + t];
      tot[t] = s;
    }
    __syncthreads();
    return;
  }
  double* buf = sy.partials + (size_t)(sy.gen & 1) * W * kPartialStride;
  ++sy.gen;
  sy.group_arrivals += (uint32_t)count * sy.group_size;
  sy.top_arrivals += sy.groups;
  if (t < count) {
    buf[sy.rank * kPartialStride + t] = wg;
    const uint32_t old = __hip_atomic_fetch_add(sy.group, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == sy.group_arrivals) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_fetch_add(sy.top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (t == 0) {
    while (__hip_atomic_load(sy.top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sy.top_arrivals)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  {
    const int pair = t & 15, grp = t >> 4;
    double s0 = 0.0, s1 = 0.0;
    if (2 * pair < count) {
      const double* col = buf + 2 * pair;
      for (int r0 = grp; r0 < W; r0 += 8 * kRankGroups) {
        double2 x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int r = r0 + k * kRankGroups;
          x[k] = r < W ? *reinterpret_cast<const double2*>(col + (size_t)r * kPartialStride) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          s0 += x[k].x;
          s1 += x[k].y;
        }
      }
    }
    lds[grp * kPartialStride + 2 * pair] = s0;
    lds[grp * kPartialStride + 2 * pair + 1] = s1;
  }
  __syncthreads();
  if (t < count) {
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < kRankGroups; ++g) s += lds[g * kPartialStride + t];
    tot[t] = s;
  }
  __syncthreads();
}

// RegularizationTerm::compute_derivatives (optimization/regularization.rs:102-159):
// score, gradient (x, y) and Hessian (xy block); zero when inactive.
struct RegTerms {
  double s, g0, g1, h00, h01, h11;
};
__device__ inline RegTerms regularization_xy(const NewtonConfig& cfg, const double* pose, uint32_t corr) {
  RegTerms r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (!cfg.reg_active) return r;
  const double scale = cfg.reg_scale, w = (double)corr;
  const double dx = cfg.reg_t[0] - pose[0], dy = cfg.reg_t[1] - pose[1];
  double sy, cy;
  sincos(pose[5], &sy, &cy);
  const double lon = dy * sy + dx * cy;
  r.s = -scale * w * lon * lon;
  r.g0 = scale * w * 2.0 * cy * lon;
  r.g1 = scale * w * 2.0 * sy * lon;
  r.h00 = -scale * w * 2.0 * cy * cy;
  r.h01 = -scale * w * 2.0 * cy * sy;
  r.h11 = -scale * w * 2.0 * sy * sy;
  return r;
}

// Element (i, j) of the regularisation Hessian.
__device__ inline double reg_hessian(const RegTerms& r, int i, int j) {
  return (i > 1 || j > 1) ? 0.0 : (i == 0 && j == 0 ? r.h00 : (i == 1 && j == 1 ? r.h11 : r.h01));
}

__device__ __attribute__((always_inline)) inline void regularization_terms(const NewtonConfig& cfg, const double* pose, uint32_t corr, double* rs,
                                     double rg[6], double rH[36]) {
  const RegTerms r = regularization_xy(cfg, pose, corr);
  *rs = r.s;
  for (int i = 0; i < 6; ++i) rg[i] = i == 0 ? r.g0 : (i == 1 ? r.g1 : 0.0);
  for (int i = 0; i < 36; ++i) rH[i] = reg_hessian(r, i / 6, i % 6);
}

__device__ uint32_t oscillation_count(const double* h, int k) {
  // count_oscillation (optimization/oscillation.rs:50-99)
  if (k < 3) return 0;
  uint32_t cnt = 0, mx = 0;
  for (int i = 2; i < k; ++i) {
    double cur[3], prev[3];
    for (int a = 0; a < 3; ++a) {
      cur[a] = h[3 * i + a] - h[3 * (i - 1) + a];
      prev[a] = h[3 * (i - 1) + a] - h[3 * (i - 2) + a];
    }
    const double cn = sqrt(cur[0] * cur[0] + cur[1] * cur[1] + cur[2] * cur[2]);
    const double pn = sqrt(prev[0] * prev[0] + prev[1] * prev[1] + prev[2] * prev[2]);
    if (cn < 1e-10 || pn < 1e-10) {
      cnt = 0;
      continue;
    }
    double c = 0.0;
    for (int a = 0; a < 3; ++a) c += (cur[a] / cn) * (prev[a] / pn);
    cnt = (c < -0.9) ? cnt + 1 : 0;
    mx = cnt > mx ? cnt : mx;
  }
  return mx;
}

// Points [begin, last) of workgroup `rank` among the W of a slot.
__device__ inline void slot_range(uint32_t n, int W, int rank, uint32_t& begin, uint32_t& last) {
  const uint32_t per_wg = (n + (uint32_t)W - 1) / (uint32_t)W;
  begin = min(n, (uint32_t)rank * per_wg);
  last = min(n, begin + per_wg);
}

// Exchange state of workgroup `rank` of slot `slot`.
__device__ inline SlotSync make_slot_sync(double* partials, uint32_t* counters, int slot, int W, int rank, int E) {
  SlotSync sy;
  sy.E = E;
  sy.partials = partials + (size_t)slot * (E > 0 ? E : 2) * W * kPartialStride;
  uint32_t* lines = counters + (size_t)slot * counter_lines(W) * kCounterStride;
  sy.top = lines;
  sy.group = lines + (size_t)(1 + rank / kArrivalGroup) * kCounterStride;
  sy.group_size = (uint32_t)min(kArrivalGroup, W - (rank / kArrivalGroup) * kArrivalGroup);
  sy.groups = (uint32_t)((W + kArrivalGroup - 1) / kArrivalGroup);
  sy.gen = 0;
  sy.group_arrivals = 0;
  sy.top_arrivals = 0;
  sy.W = W;
  sy.rank = rank;
  return sy;
}

// One whole alignment by the W workgroups of a slot; this workgroup is
// sy.rank.  Workgroup 0 writes *out.  kF32: per-point terms in FP32 (the
// precision variant); FP64 otherwise.  The exchange state sy carries over
// from one alignment to the next (the persistent server runs many).
template <bool kF32>
__device__ __attribute__((always_inline)) void align_slot(const GridView& g, const NewtonConfig& cfg,
                                                        const SlotJob& job, ndt_result* out, SlotSync& sy) {
  __shared__ PoseTables s_tab;
  __shared__ SlotState S;
  __shared__ SolveScratch s_solve;
  __shared__ double s_scratch[kReduceLds];
  // The derivative passes stage neighbours here; the workgroup reduction
  // reuses the same space after them.
  __shared__ union {
    float4 rec[kStageK * 3 * kBlock];
    double red[(kAccVals + 1) * (kBlock + 8)];
  } s_big;
  __shared__ int32_t s_stage_slot[kStageK * kBlock];
  __shared__ double s_tot[kAccSlots];
  double* const s_red = s_big.red;
  const Stage stage{s_big.rec, s_stage_slot};
  __shared__ double s_q[4];
  __shared__ double s_tp[6], s_trig[13];  // trial pose; sin/cos values of the pose tables

  const int W = sy.W;
  const int rank = sy.rank;
  // Workgroup `rank` owns points [begin, last); its thread t takes local
  // index (t % 64) * 4 + t / 64 and then steps by the block size, so a short
  // range is shared by all four waves (more loads in flight per CU).
  uint32_t begin, last;
  slot_range(job.n, W, rank, begin, last);
  const uint32_t first = begin + (threadIdx.x % kWave) * (kBlock / kWave) + threadIdx.x / kWave;
  const uint32_t stride = kBlock;
  const int max_iter = cfg.max_iter < (uint32_t)kMaxIter ? (int)cfg.max_iter : kMaxIter;
  // Phase stamps: workgroup 0 thread 0 only.
  const bool rec = cfg.phase != nullptr && blockIdx.x == 0 && threadIdx.x == 0;
  // Every workgroup's thread 0 times its own derivative passes (for the
  // slowest-workgroup figure).
  const bool rec_all = cfg.phase != nullptr && threadIdx.x == 0;
  unsigned long long own_deriv = 0, t_own = 0, t_solve0 = 0, t_solve1 = 0, ph_pre = 0, ph_step = 0;
  unsigned long long ph[6] = {0, 0, 0, 0, 0, 0}, t_iter = 0, t_blk = 0, t_red = 0, t_start = 0;
  unsigned long long n_ls = 0, n_it = 0;
  if (rec) t_start = wall_clock64();

  if (threadIdx.x == 0) {
    for (int i = 0; i < 6; ++i) {
      S.pose[i] = job.pose[i];
      S.best_pose[i] = job.pose[i];
    }
    for (int i = 0; i < 36; ++i) S.last_H[i] = 0.0;
    S.best_score = -HUGE_VAL;
    S.hist[0] = job.pose[0];
    S.hist[1] = job.pose[1];
    S.hist[2] = job.pose[2];
    S.nhist = 1;
    S.status = -1;
    S.last_corr = 0;
    S.corr_total = 0;
    s_solve.sweeps = 0;
    s_solve.solves = 0;
    newton_step_reset(s_solve);
  }
  // Pose tables: wave 0 (three lanes evaluate the angles' sin/cos at once).
  if (threadIdx.x < kWave) make_pose_tables_wave(S.pose, s_trig, &s_tab);
  __syncthreads();

  for (int it = 0; it < max_iter; ++it) {
    if (rec) t_iter = wall_clock64();
    if (rec_all) t_own = wall_clock64();
    double acc[kAccSlots];
#pragma unroll
    for (int i = 0; i < kAccSlots; ++i) acc[i] = 0.0;
    uint32_t corr = 0;
    derivative_points<true, typename std::conditional<kF32, float, double>::type>(g, cfg, s_tab, job.pts, last,
                                                                                   first, stride, acc, corr, stage);
    acc[kAccVals] = (double)corr;
    slot_reduce(acc, kAccVals + 1, s_scratch, s_red, s_tot, sy, rec_all ? &t_blk : nullptr);
    if (rec_all) own_deriv += t_blk - t_own;
    if (rec) {
      t_red = wall_clock64();
      ph[0] += t_blk - t_iter;
      ph[1] += t_red - t_blk;
      ++n_it;
    }

    if (threadIdx.x == 0) {
      const uint32_t c = (uint32_t)s_tot[kAccVals];
      S.corr_total += c;
      S.ctrl = 0;
      S.need_solve = 0;
      if (c < 10) {  // solver.rs:168-174
        if (it == 0) {
          S.status = NDT_NO_CORRESPONDENCES;
          S.ctrl = 1;
        } else {
          S.ctrl = 3;  // break: max-iterations result
        }
      } else {
        S.last_corr = c;
        S.score = s_tot[0] + regularization_xy(cfg, S.pose, c).s;
        if (S.score > S.best_score) {
          S.best_score = S.score;
          for (int i = 0; i < 6; ++i) S.best_pose[i] = S.pose[i];
        }
        S.need_solve = 1;
        newton_step_reset(s_solve);
      }
    }
    // Gradient and full Hessian (plus the regularisation terms, which touch
    // only g0, g1 and the xy block) written by the lanes of wave 0 at once.
    if (threadIdx.x < kWave && (uint32_t)s_tot[kAccVals] >= 10) {
      const int l = threadIdx.x;
      const RegTerms r = regularization_xy(cfg, S.pose, (uint32_t)s_tot[kAccVals]);
      if (l < 36) {
        const int i = l / 6, j = l % 6;
        S.last_H[l] = s_tot[7 + (i >= j ? tri(i, j) : tri(j, i))] + reg_hessian(r, i, j);
      }
      if (l < 6) S.grad[l] = s_tot[1 + l] + (l == 0 ? r.g0 : (l == 1 ? r.g1 : 0.0));
    }
    __syncthreads();
    if (rec) t_solve0 = wall_clock64();
    if (S.need_solve) newton_step_wg(S.last_H, S.grad, S.delta, s_solve);
    if (rec) {
      t_solve1 = wall_clock64();
      ph_pre += t_solve0 - t_red;
      ph_step += t_solve1 - t_solve0;
    }
    if (threadIdx.x == 0 && S.need_solve) {
      const uint32_t c = S.last_corr;
      if (!s_solve.ok) {  // solver.rs:204-227
        S.status = NDT_SINGULAR_HESSIAN;
        for (int i = 0; i < 6; ++i) S.final_pose[i] = S.best_pose[i];
        S.final_score = S.best_score;
        S.iterations = it;
        for (int i = 0; i < 36; ++i) S.final_H[i] = S.last_H[i];
        S.final_corr = S.last_corr;
        S.ctrl = 1;
      } else {
        double norm = 0.0;
        for (int i = 0; i < 6; ++i) norm += S.delta[i] * S.delta[i];
        norm = sqrt(norm);
        if (norm < cfg.trans_eps) {  // solver.rs:229-253
          S.status = NDT_CONVERGED;
          for (int i = 0; i < 6; ++i) S.final_pose[i] = S.pose[i];
          S.final_score = S.score;
          S.iterations = it + 1;
          for (int i = 0; i < 36; ++i) S.final_H[i] = S.last_H[i];
          S.final_corr = c;
          S.ctrl = 1;
        } else {
          double gd = 0.0;
          for (int i = 0; i < 6; ++i) {
            S.dir[i] = S.delta[i] / norm;
            gd += S.grad[i] * S.dir[i];
          }
          if (gd <= 0.0)
            for (int i = 0; i < 6; ++i) S.dir[i] = -S.dir[i];
          if (cfg.use_ls) {
            double d0 = 0.0;
            for (int i = 0; i < 6; ++i) d0 += S.grad[i] * S.dir[i];
            if (d0 <= 0.0) {
              S.step = cfg.step_size;
            } else {
              mt_begin(S.mt, S.score, d0, cfg.step_size);
              S.ctrl = 2;
            }
          } else {
            S.step = fmin(norm, cfg.step_size);
          }
        }
      }
      if (S.ctrl == 2)
        for (int i = 0; i < 6; ++i) s_tp[i] = S.pose[i] + S.dir[i] * S.mt.a_t;
    }
    if (threadIdx.x < kWave) {
      wave_lds_fence();
      if (S.ctrl == 2) make_pose_tables_wave(s_tp, s_trig, &s_tab);
    }
    __syncthreads();
    if (rec) ph[2] += wall_clock64() - t_red;
    const int ctrl = S.ctrl;
    if (ctrl == 1 || ctrl == 3) break;
    if (ctrl == 2) {
      const unsigned long long t_ls = rec ? wall_clock64() : 0ull;
      // line_search_with_regularization (solver.rs:640-699): phi(a) = score and
      // directional derivative at pose + a*dir, without the Hessian.
      for (;;) {
        double la[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) la[i] = 0.0;
        uint32_t lc = 0;
        derivative_points<false, typename std::conditional<kF32, float, double>::type>(
            g, cfg, s_tab, job.pts, last, first, stride, la, lc, stage);
        la[7] = (double)lc;
        slot_reduce(la, 8, s_scratch, s_red, s_tot, sy);
        if (threadIdx.x == 0) {
          double tp[6];
          for (int i = 0; i < 6; ++i) tp[i] = S.pose[i] + S.dir[i] * S.mt.a_t;
          double rs, rg[6], rH[36];
          regularization_terms(cfg, tp, (uint32_t)s_tot[7], &rs, rg, rH);
          double dd = 0.0;
          for (int i = 0; i < 6; ++i) dd += (s_tot[1 + i] + rg[i]) * S.dir[i];
          mt_feed(S.mt, s_tot[0] + rs, dd);
          if (S.mt.done) {
            S.step = S.mt.conv ? S.mt.a_t : kMtMax;
          } else {
            for (int i = 0; i < 6; ++i) s_tp[i] = S.pose[i] + S.dir[i] * S.mt.a_t;
          }
        }
        if (threadIdx.x < kWave) {
          wave_lds_fence();
          if (!S.mt.done) make_pose_tables_wave(s_tp, s_trig, &s_tab);
        }
        __syncthreads();
        if (rec) ++n_ls;
        if (S.mt.done) break;
      }
      if (rec) ph[3] += wall_clock64() - t_ls;
    }
    if (threadIdx.x == 0) {
      for (int i = 0; i < 6; ++i) S.pose[i] = S.pose[i] + S.dir[i] * S.step;  // types.rs:183-197
      S.hist[3 * S.nhist + 0] = S.pose[0];
      S.hist[3 * S.nhist + 1] = S.pose[1];
      S.hist[3 * S.nhist + 2] = S.pose[2];
      S.nhist += 1;
    }
    if (threadIdx.x < kWave) make_pose_tables_wave(S.pose, s_trig, &s_tab);
    __syncthreads();
  }

  if (threadIdx.x == 0 && S.status < 0) {  // max iterations (solver.rs:286-304)
    S.status = NDT_MAX_ITERATIONS;
    for (int i = 0; i < 6; ++i) S.final_pose[i] = S.best_pose[i];
    S.final_score = S.best_score;
    S.iterations = (int)cfg.max_iter;
    for (int i = 0; i < 36; ++i) S.final_H[i] = S.last_H[i];
    S.final_corr = S.last_corr;
  }
  if (threadIdx.x == 0) quat_from_euler(S.final_pose[3], S.final_pose[4], S.final_pose[5], s_q);
  __syncthreads();

  const unsigned long long t_nv = rec ? wall_clock64() : 0ull;
  double nv_sum = 0.0, nv_cnt = 0.0;
  const bool no_corr = S.status == NDT_NO_CORRESPONDENCES;
  if (!no_corr) nvtl_points(g, cfg.d1, cfg.d2, s_q, S.final_pose, job.pts, last, first, stride, nv_sum, nv_cnt);
  double nv[2] = {nv_sum, nv_cnt};
  slot_reduce(nv, 2, s_scratch, s_red, s_tot, sy);
  if (rec) {
    const unsigned long long t_end = wall_clock64();
    ph[4] = t_end - t_nv;
    ph[5] = t_end - t_start;
    for (int i = 0; i < 6; ++i) cfg.phase[i] = ph[i];
    cfg.phase[6] = n_it;
    cfg.phase[7] = n_ls;
    cfg.phase[9] = ph_pre;
    cfg.phase[10] = ph_step;
    cfg.phase[11] = (unsigned long long)s_solve.sweeps;
    cfg.phase[12] = (unsigned long long)s_solve.solves;
  }
  if (rec_all) atomicMax(&cfg.phase[8], own_deriv);

  if (rank == 0 && threadIdx.x == 0) {
    ndt_result r;
    if (no_corr) {  // NdtResult::no_correspondences (optimization/types.rs:133-145)
      for (int i = 0; i < 6; ++i) r.pose[i] = job.pose[i];
      r.status = NDT_NO_CORRESPONDENCES;
      r.converged = 0;
      r.score = 0.0;
      r.transform_probability = 0.0;
      r.nvtl = 0.0;
      r.iterations = 0;
      r.num_correspondences = 0;
      r.oscillation_count = 0;
      for (int i = 0; i < 36; ++i) r.hessian[i] = 0.0;
    } else {
      for (int i = 0; i < 6; ++i) r.pose[i] = S.final_pose[i];
      r.status = S.status;
      r.converged = S.status == NDT_CONVERGED;
      r.score = S.final_score;
      r.transform_probability = job.n > 0 ? S.final_score / (double)job.n : 0.0;
      r.nvtl = s_tot[1] > 0.0 ? s_tot[0] / s_tot[1] : 0.0;
      r.iterations = (uint32_t)S.iterations;
      r.num_correspondences = S.final_corr;
      r.oscillation_count = oscillation_count(S.hist, S.nhist);
      for (int i = 0; i < 36; ++i) r.hessian[i] = S.final_H[i];
    }
    r.pair_evaluations = S.corr_total;
    *out = r;
  }
}

template <bool kF32>
__global__ __launch_bounds__(kBlock) void newton_kernel(GridView g, NewtonConfig cfg, const SlotJob* __restrict__ jobs,
                                                        ndt_result* __restrict__ out, double* __restrict__ partials,
                                                        uint32_t* __restrict__ counters) {
  const int W = cfg.wg_per_slot;
  const int slot = blockIdx.x / W;
  SlotSync sy = make_slot_sync(partials, counters, slot, W, blockIdx.x % W, cfg.exchanges);
  const SlotJob job = jobs[slot];
  align_slot<kF32>(g, cfg, job, out + slot, sy);
}

// ---------------------------------------------------------------------------
// Persistent server: each slot's workgroups run the jobs posted to the slot's
// mailbox one after another (continuous batching: a slot that finishes takes
// the next request while the others keep running).
//
// Workgroup 0 of a slot polls the host mailbox and announces its decision in
// the slot's command word in device memory: the job number to run, or
// kExitCmd.  The other workgroups of the slot follow that word, so a slot
// never splits between running a job and leaving.  Leaving is launch-wide
// (ServerControl): on stop, when another slot has set exited, or after
// idle_ticks of the wall clock in which neither the host's heartbeat nor the
// launch's count of finished jobs moved (a slot that serves keeps the idle
// ones alive).
constexpr unsigned long long kExitCmd = ~0ull;

// The next command of a slot's workgroup 0: done + 1 when a job is posted,
// kExitCmd when the launch ends.  Bounded: the heartbeat timeout ends every
// wait once the host stops calling.
__device__ unsigned long long server_next_command(const ServerSlot* mb, ServerControl* ctl, unsigned long long done,
                                                  int slot, const unsigned long long* activity) {
  unsigned long long hb = __hip_atomic_load(&ctl->heartbeat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  unsigned long long act = __hip_atomic_load(activity, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long t0 = wall_clock64();
  for (;;) {
    if (__hip_atomic_load(&mb->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) > done) return done + 1;
    if (__hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) ||
        __hip_atomic_load(&ctl->exited, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
      return kExitCmd;
    const unsigned long long h = __hip_atomic_load(&ctl->heartbeat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const unsigned long long a = __hip_atomic_load(activity, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long now = wall_clock64();
    if (h != hb || a != act) {  // a host call, or another slot finished a job
      hb = h;
      act = a;
      t0 = now;
    } else {
      const unsigned long long limit = __hip_atomic_load(&ctl->idle_ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (now - t0 > limit) {
        __hip_atomic_store(&ctl->exit_slot, (unsigned long long)slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&ctl->exit_waited, now - t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&ctl->exit_timeout, limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&ctl->exit_heartbeat, hb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&ctl->exited, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        return kExitCmd;
      }
    }
    __builtin_amdgcn_s_sleep(127);
  }
}

// The host leaves each request's points in pinned memory (host_points); at
// the start of a job every workgroup copies its own point range into the
// slot's device buffer, which is all it reads afterwards.  No runtime copy
// is needed while the launch occupies the device.
template <bool kF32>
__global__ __launch_bounds__(kBlock) void server_kernel(GridView g, NewtonConfig cfg, ServerSlot* __restrict__ mb,
                                                        ServerControl* __restrict__ ctl,
                                                        const float* __restrict__ host_points,
                                                        float* __restrict__ slot_points, uint32_t max_points,
                                                        double* __restrict__ partials, uint32_t* __restrict__ counters,
                                                        unsigned long long* __restrict__ cmd) {
  const int W = cfg.wg_per_slot;
  const int slot = blockIdx.x / W;
  const int rank = blockIdx.x % W;
  SlotSync sy = make_slot_sync(partials, counters, slot, W, rank, 0);
  unsigned long long* my_cmd = cmd + (size_t)slot * kCounterStride / 2;
  // One line after the command words: jobs finished by the launch.
  unsigned long long* activity = cmd + (size_t)gridDim.x / W * kCounterStride / 2;
  __shared__ unsigned long long s_cmd;
  __shared__ SlotJob s_job;
  unsigned long long done = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      unsigned long long c;
      if (rank == 0) {
        c = server_next_command(&mb[slot], ctl, done, slot, activity);
        // Acquire the host's writes (points, n, pose, released with seq)
        // before publishing the command with release semantics, so the
        // followers' acquire of the command orders their reads after them.
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        __hip_atomic_store(my_cmd, c, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        // Workgroup 0 writes the command word within its bounded wait.
        for (;;) {
          c = __hip_atomic_load(my_cmd, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          if (c > done) break;
          __builtin_amdgcn_s_sleep(8);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
      }
      if (c != kExitCmd) {
        s_job.pts = slot_points + (size_t)slot * max_points * 3;
        s_job.n = mb[slot].n;
        s_job.pad = 0;
        for (int i = 0; i < 6; ++i) s_job.pose[i] = mb[slot].pose[i];
      }
      s_cmd = c;
    }
    __syncthreads();
    if (s_cmd == kExitCmd) break;
    const SlotJob job = s_job;
    {
      uint32_t begin, last;
      slot_range(job.n, W, rank, begin, last);
      const float* src = host_points + (size_t)slot * max_points * 3;
      float* dst = slot_points + (size_t)slot * max_points * 3;
      for (uint32_t k = 3 * begin + threadIdx.x; k < 3 * last; k += kBlock) dst[k] = src[k];
    }
    __syncthreads();
    align_slot<kF32>(g, cfg, job, &mb[slot].result, sy);
    ++done;
    if (rank == 0 && threadIdx.x == 0) {
      __hip_atomic_store(&mb[slot].done, done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_fetch_add(activity, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void store_job_kernel(SlotJob job, SlotJob* __restrict__ dst) { *dst = job; }

__global__ void fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void zero_u32(uint32_t* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Single derivative pass (parity entry point).
template <bool kF32>
__global__ __launch_bounds__(kBlock) void derivative_kernel(GridView g, NewtonConfig cfg, const float* __restrict__ pts,
                                                            uint32_t n, PoseTables T, int hess,
                                                            double* __restrict__ partials) {
  __shared__ double s_scratch[(kBlock / kWave) * kAccSlots];
  __shared__ float4 s_stage_rec[kStageK * 3 * kBlock];
  __shared__ int32_t s_stage_slot[kStageK * kBlock];
  const Stage stage{s_stage_rec, s_stage_slot};
  double acc[kAccSlots];
#pragma unroll
  for (int i = 0; i < kAccSlots; ++i) acc[i] = 0.0;
  uint32_t corr = 0;
  const uint32_t first = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
  using R = typename std::conditional<kF32, float, double>::type;
  if (hess)
    derivative_points<true, R>(g, cfg, T, pts, n, first, stride, acc, corr, stage);
  else
    derivative_points<false, R>(g, cfg, T, pts, n, first, stride, acc, corr, stage);
  acc[kAccVals] = (double)corr;
  const double tot = block_sums<kBlock, kAccSlots>(acc, kAccVals + 1, s_scratch);
  if (threadIdx.x < kAccVals + 1) partials[blockIdx.x * kPartialStride + threadIdx.x] = tot;
}

__global__ void reduce_partials(const double* __restrict__ partials, int blocks, int count, double* __restrict__ out) {
  // out[pose * count + v] = sum over blocks, fixed order; grid.y = pose.
  const int v = threadIdx.x;
  if (v >= count) return;
  const double* p = partials + (size_t)blockIdx.y * blocks * kPartialStride;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += p[b * kPartialStride + v];
  out[blockIdx.y * count + v] = s;
}

// Scoring: kind 0 NVTL (radius, max), 1 TP (radius, sum over pairs),
// 2 per-point sums, 3 NVTL with the 3x3x3 cell scan.  grid = (blocks, poses).
__global__ __launch_bounds__(kBlock) void score_kernel(GridView g, double d1, double d2, int kind,
                                                       const float* __restrict__ pts, uint32_t n,
                                                       const double* __restrict__ poses, double* __restrict__ partials,
                                                       double* __restrict__ per_point) {
  __shared__ double s_scratch[(kBlock / kWave) * 2];
  __shared__ double s_q[4];
  __shared__ double s_t[3];
  const double* pose = poses + 6 * blockIdx.y;
  if (threadIdx.x == 0) {
    quat_from_euler(pose[3], pose[4], pose[5], s_q);
    s_t[0] = pose[0];
    s_t[1] = pose[1];
    s_t[2] = pose[2];
  }
  __syncthreads();
  double sum = 0.0, cnt = 0.0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double x[3];
    iso_transform(s_q, s_t, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], x);
    if (kind == kScoreNvtl || kind == kScoreNvtlCell27) {
      double best = -HUGE_VAL;
      bool any = false;
      auto vis = [&](int32_t, VoxelRecord r) {
        best = fmax(best, point_score(x, r.m0, r.m1, r.m2, d1, d2));
        any = true;
      };
      if (kind == kScoreNvtl)
        for_each_radius<ProbeNarrow>(g, (float)x[0], (float)x[1], (float)x[2], vis);
      else
        for_each_cell27<ProbeNarrow>(g, x[0], x[1], x[2], vis);
      if (any) {
        sum += best;
        cnt += 1.0;
      }
    } else {
      double ps = 0.0;
      double pc = 0.0;
      for_each_radius<ProbeNarrow>(g, (float)x[0], (float)x[1], (float)x[2], [&](int32_t, VoxelRecord r) {
        ps += point_score(x, r.m0, r.m1, r.m2, d1, d2);
        pc += 1.0;
      });
      if (kind == kScorePerPoint) {
        per_point[i] = ps;
      } else {
        sum += ps;
        cnt += pc;
      }
    }
  }
  if (kind == kScorePerPoint) return;
  double v[2] = {sum, cnt};
  const double tot = block_sums<kBlock, 2>(v, 2, s_scratch);
  if (threadIdx.x < 2)
    partials[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kPartialStride + threadIdx.x] = tot;
}

__global__ void finish_scores(const double* __restrict__ sums, int m, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double s = sums[2 * i], c = sums[2 * i + 1];
  out[i] = c > 0.0 ? s / c : 0.0;
}

__global__ __launch_bounds__(kBlock) void lookup_kernel(GridView g, int mode, const float* __restrict__ q, uint32_t n,
                                                        uint32_t max_k, uint32_t* __restrict__ counts,
                                                        int32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t c = 0;
  auto vis = [&](int32_t s, VoxelRecord) {
    if (c < max_k) idx[(size_t)i * max_k + c] = s;
    ++c;
  };
  if (mode == 4)
    for_each_cell27<ProbeNarrow>(g, (double)q[3 * i], (double)q[3 * i + 1], (double)q[3 * i + 2], vis);
  else
    for_each_neighbor<ProbeNarrow>(g, mode, q[3 * i], q[3 * i + 1], q[3 * i + 2], vis);
  for (uint32_t k = c; k < max_k; ++k) idx[(size_t)i * max_k + k] = -1;
  counts[i] = c;
}

}  // namespace

// Workspace layout: partials (slots x (E or 2) x W x kPartialStride doubles),
// then the arrival counters of the alternating protocol.
constexpr size_t partials_bytes(int slots, int wg_per_slot, int exchanges) {
  return (size_t)slots * (exchanges > 0 ? exchanges : 2) * wg_per_slot * kPartialStride * sizeof(double);
}
constexpr size_t workspace_bytes(int slots, int wg_per_slot, int exchanges) {
  return partials_bytes(slots, wg_per_slot, exchanges) +
         (size_t)slots * counter_lines(wg_per_slot) * kCounterStride * sizeof(uint32_t) + 256;
}
static_assert(workspace_bytes(1, kNoAllocMaxWorkgroups, 0) == kNoAllocWorkspaceBytes, "allocation-free workspace size");
static_assert(sizeof(SlotJob) <= kNoAllocWorkspaceOffset, "allocation-free job slot");

size_t newton_workspace_bytes(int slots, int wg_per_slot, int exchanges) {
  return workspace_bytes(slots, wg_per_slot, exchanges);
}

hipError_t store_job(const SlotJob& job, SlotJob* d_job, hipStream_t st) {
  hipLaunchKernelGGL(store_job_kernel, dim3(1), dim3(1), 0, st, job, d_job);
  return hipGetLastError();
}

int newton_max_coresident_workgroups() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  // The smallest over the persistent kernels (both precisions, one-shot and
  // server), so any of them fits a launch sized by this.
  const void* fns[] = {reinterpret_cast<const void*>(newton_kernel<false>),
                       reinterpret_cast<const void*>(newton_kernel<true>),
                       reinterpret_cast<const void*>(server_kernel<false>),
                       reinterpret_cast<const void*>(server_kernel<true>)};
  per_cu = 1 << 30;
  for (const void* fn : fns) {
    int k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, fn, kBlock, 0) != hipSuccess) return 0;
    per_cu = std::min(per_cu, k);
  }
  return cus * per_cu;
}

hipError_t launch_newton(const GridView& g, const NewtonConfig& cfg, const SlotJob* d_jobs, int slots,
                         ndt_result* d_out, void* ws, hipStream_t st, bool cooperative, hipEvent_t ev_start) {
  const int W = cfg.wg_per_slot;
  double* partials = reinterpret_cast<double*>(ws);
  uint32_t* counters =
      reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + partials_bytes(slots, W, cfg.exchanges));
  if (cfg.exchanges > 0) {
    const size_t nfill = partials_bytes(slots, W, cfg.exchanges) / sizeof(double);
    hipLaunchKernelGGL(fill_u64, dim3((unsigned)std::min<size_t>(2048, (nfill + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long*>(partials), nfill, kExchangeEmpty);
  } else {
    const int ncounters = slots * counter_lines(W) * kCounterStride;
    hipLaunchKernelGGL(zero_u32, dim3(std::min(1024, (ncounters + 255) / 256)), dim3(256), 0, st, counters, ncounters);
  }
  if (ev_start) {
    const hipError_t e = hipEventRecord(ev_start, st);
    if (e != hipSuccess) return e;
  }
  const dim3 grid((unsigned)(slots * W)), block(kBlock);
  if (cooperative) {
    GridView gg = g;
    NewtonConfig cc = cfg;
    const SlotJob* jj = d_jobs;
    ndt_result* oo = d_out;
    void* args[] = {&gg, &cc, &jj, &oo, &partials, &counters};
    const void* fn = cfg.precision == NDT_PRECISION_FP32 ? reinterpret_cast<const void*>(newton_kernel<true>)
                                                         : reinterpret_cast<const void*>(newton_kernel<false>);
    return hipLaunchCooperativeKernel(fn, grid, block, args, 0, st);
  }
  if (cfg.precision == NDT_PRECISION_FP32)
    hipLaunchKernelGGL(newton_kernel<true>, grid, block, 0, st, g, cfg, d_jobs, d_out, partials, counters);
  else
    hipLaunchKernelGGL(newton_kernel<false>, grid, block, 0, st, g, cfg, d_jobs, d_out, partials, counters);
  return hipGetLastError();
}

// The same work as launch_newton, recorded once as a graph: jobs copy from
// pinned host memory, counter fill (a kernel: no memset nodes), start event,
// the persistent Newton kernel (cooperative attribute when W > 1), end event,
// results copy to pinned host memory.  Replays read whatever h_jobs holds at
// launch time, so a caller refills h_jobs and replays.
hipError_t build_newton_graph(const GridView& g, const NewtonConfig& cfg, SlotJob* d_jobs, const SlotJob* h_jobs,
                              int slots, ndt_result* d_out, ndt_result* h_out, void* ws, hipEvent_t ev0,
                              hipEvent_t ev1, hipGraph_t* graph_out, hipGraphExec_t* exec_out) {
  *graph_out = nullptr;
  *exec_out = nullptr;
  const int W = cfg.wg_per_slot;
  double* partials = reinterpret_cast<double*>(ws);
  uint32_t* counters =
      reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + partials_bytes(slots, W, cfg.exchanges));
  int ncounters = slots * counter_lines(W) * kCounterStride;
  unsigned long long* fill_ptr = reinterpret_cast<unsigned long long*>(partials);
  size_t nfill = partials_bytes(slots, W, cfg.exchanges) / sizeof(double);
  unsigned long long fill_val = kExchangeEmpty;
  hipGraph_t graph = nullptr;
  hipError_t e = hipGraphCreate(&graph, 0);
  if (e != hipSuccess) return e;
  hipGraphNode_t n_jobs, n_zero, n_ev0, n_newton, n_ev1, n_res;
  e = hipGraphAddMemcpyNode1D(&n_jobs, graph, nullptr, 0, d_jobs, h_jobs, (size_t)slots * sizeof(SlotJob),
                              hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    // Exchange buffers to the sentinel (write-once protocol) or counters to
    // zero, by a kernel node: no memset nodes.
    hipKernelNodeParams kp{};
    void* zargs[] = {&counters, &ncounters};
    void* fargs[] = {&fill_ptr, &nfill, &fill_val};
    if (cfg.exchanges > 0) {
      kp.func = reinterpret_cast<void*>(fill_u64);
      kp.gridDim = dim3((unsigned)std::min<size_t>(2048, (nfill + 255) / 256));
      kp.kernelParams = fargs;
    } else {
      kp.func = reinterpret_cast<void*>(zero_u32);
      kp.gridDim = dim3(std::min(1024, (ncounters + 255) / 256));
      kp.kernelParams = zargs;
    }
    kp.blockDim = dim3(256);
    kp.sharedMemBytes = 0;
    kp.extra = nullptr;
    e = hipGraphAddKernelNode(&n_zero, graph, &n_jobs, 1, &kp);
  }
  if (e == hipSuccess) e = hipGraphAddEventRecordNode(&n_ev0, graph, &n_zero, 1, ev0);
  GridView gg = g;
  NewtonConfig cc = cfg;
  const SlotJob* jj = d_jobs;
  ndt_result* oo = d_out;
  void* args[] = {&gg, &cc, &jj, &oo, &partials, &counters};
  if (e == hipSuccess) {
    hipKernelNodeParams kp{};
    kp.func = cfg.precision == NDT_PRECISION_FP32 ? reinterpret_cast<void*>(newton_kernel<true>)
                                                  : reinterpret_cast<void*>(newton_kernel<false>);
    kp.gridDim = dim3((unsigned)(slots * W));
    kp.blockDim = dim3(kBlock);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    e = hipGraphAddKernelNode(&n_newton, graph, &n_ev0, 1, &kp);
  }
  if (e == hipSuccess && W > 1) {
    hipKernelNodeAttrValue v{};
    v.cooperative = 1;
    e = hipGraphKernelNodeSetAttribute(n_newton, hipKernelNodeAttributeCooperative, &v);
  }
  if (e == hipSuccess) e = hipGraphAddEventRecordNode(&n_ev1, graph, &n_newton, 1, ev1);
  if (e == hipSuccess)
    e = hipGraphAddMemcpyNode1D(&n_res, graph, &n_ev1, 1, h_out, d_out, (size_t)slots * sizeof(ndt_result),
                                hipMemcpyDeviceToHost);
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return e;
  }
  *graph_out = graph;
  *exec_out = exec;
  return hipSuccess;
}

size_t server_workspace_bytes(int slots, int wg_per_slot) {
  // counters, then a command word line per slot, then the activity line
  return newton_workspace_bytes(slots, wg_per_slot) + (size_t)(slots + 1) * kCounterStride * sizeof(uint32_t);
}

hipError_t launch_server(const GridView& g, const NewtonConfig& cfg, ServerSlot* mailbox, ServerControl* ctl,
                         int slots, const float* host_points, float* slot_points, uint32_t max_points, void* ws,
                         hipStream_t st) {
  const int W = cfg.wg_per_slot;
  double* partials = reinterpret_cast<double*>(ws);
  uint32_t* counters =
      reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + (size_t)slots * 2 * W * kPartialStride * sizeof(double));
  const int nc = slots * counter_lines(W) * kCounterStride;
  unsigned long long* cmd = reinterpret_cast<unsigned long long*>(counters + nc);
  const int nall = nc + (slots + 1) * kCounterStride;  // counters, command words and the activity line
  hipLaunchKernelGGL(zero_u32, dim3(std::min(1024, (nall + 255) / 256)), dim3(256), 0, st, counters, nall);
  GridView gg = g;
  NewtonConfig cc = cfg;
  const float* hp = host_points;
  float* pp = slot_points;
  void* args[] = {&gg, &cc, &mailbox, &ctl, &hp, &pp, &max_points, &partials, &counters, &cmd};
  const void* fn = cfg.precision == NDT_PRECISION_FP32 ? reinterpret_cast<const void*>(server_kernel<true>)
                                                       : reinterpret_cast<const void*>(server_kernel<false>);
  return hipLaunchCooperativeKernel(fn, dim3((unsigned)(slots * W)), dim3(kBlock), args, 0, st);
}

hipError_t launch_derivatives(const GridView& g, const NewtonConfig& cfg, const float* d_pts, uint32_t n,
                              const double pose[6], int hess, double* d_partials, int blocks, double* d_out,
                              hipStream_t st) {
  PoseTables T;
  // The pose tables on the host: the formulas behind kPoseTable, in host code.
  {
    const double sr = sin(pose[3]), cr = cos(pose[3]), sp = sin(pose[4]), cp = cos(pose[4]), sy = sin(pose[5]),
                 cy = cos(pose[5]);
    T.t[0] = pose[0];
    T.t[1] = pose[1];
    T.t[2] = pose[2];
    T.R[0] = cp * cy;                T.R[1] = -cp * sy;               T.R[2] = sp;
    T.R[3] = sr * sp * cy + cr * sy; T.R[4] = cr * cy - sr * sp * sy; T.R[5] = -sr * cp;
    T.R[6] = sr * sy - cr * sp * cy; T.R[7] = cr * sp * sy + sr * cy; T.R[8] = cr * cp;
    auto snap = [](double a, double* s, double* c) {
      if (fabs(a) < 1e-5) {
        *s = 0.0;
        *c = 1.0;
      } else {
        *s = sin(a);
        *c = cos(a);
      }
    };
    double sx, cx, sy2, cy2, sz, cz;
    snap(pose[3], &sx, &cx);
    snap(pose[4], &sy2, &cy2);
    snap(pose[5], &sz, &cz);
    const double s_y = sy2, c_y = cy2;
    auto& J = T.j;
    J[0][0] = -sx * sz + cx * s_y * cz; J[0][1] = -sx * cz - cx * s_y * sz; J[0][2] = -cx * c_y;
    J[1][0] = cx * sz + sx * s_y * cz;  J[1][1] = cx * cz - sx * s_y * sz;  J[1][2] = -sx * c_y;
    J[2][0] = -s_y * cz;                J[2][1] = s_y * sz;                 J[2][2] = c_y;
    J[3][0] = sx * c_y * cz;            J[3][1] = -sx * c_y * sz;           J[3][2] = sx * s_y;
    J[4][0] = -cx * c_y * cz;           J[4][1] = cx * c_y * sz;            J[4][2] = -cx * s_y;
    J[5][0] = -c_y * sz;                J[5][1] = -c_y * cz;                J[5][2] = 0.0;
    J[6][0] = cx * cz - sx * s_y * sz;  J[6][1] = -cx * sz - sx * s_y * cz; J[6][2] = 0.0;
    J[7][0] = sx * cz + cx * s_y * sz;  J[7][1] = cx * s_y * cz - sx * sz;  J[7][2] = 0.0;
    auto& H = T.h;
    H[0][0] = -cx * sz - sx * s_y * cz;  H[0][1] = -cx * cz + sx * s_y * sz;  H[0][2] = sx * c_y;
    H[1][0] = -sx * sz + cx * s_y * cz;  H[1][1] = -cx * s_y * sz - sx * cz;  H[1][2] = -cx * c_y;
    H[2][0] = cx * c_y * cz;             H[2][1] = -cx * c_y * sz;            H[2][2] = cx * s_y;
    H[3][0] = sx * c_y * cz;             H[3][1] = -sx * c_y * sz;            H[3][2] = sx * s_y;
    H[4][0] = -sx * cz - cx * s_y * sz;  H[4][1] = sx * sz - cx * s_y * cz;   H[4][2] = 0.0;
    H[5][0] = cx * cz - sx * s_y * sz;   H[5][1] = -sx * s_y * cz - cx * sz;  H[5][2] = 0.0;
    H[6][0] = -c_y * cz;                 H[6][1] = c_y * sz;                  H[6][2] = s_y;
    H[7][0] = -sx * s_y * cz;            H[7][1] = sx * s_y * sz;             H[7][2] = sx * c_y;
    H[8][0] = cx * s_y * cz;             H[8][1] = -cx * s_y * sz;            H[8][2] = -cx * c_y;
    H[9][0] = s_y * sz;                  H[9][1] = s_y * cz;                  H[9][2] = 0.0;
    H[10][0] = -sx * c_y * sz;           H[10][1] = -sx * c_y * cz;           H[10][2] = 0.0;
    H[11][0] = cx * c_y * sz;            H[11][1] = cx * c_y * cz;            H[11][2] = 0.0;
    H[12][0] = -c_y * cz;                H[12][1] = c_y * sz;                 H[12][2] = 0.0;
    H[13][0] = -cx * sz - sx * s_y * cz; H[13][1] = -cx * cz + sx * s_y * sz; H[13][2] = 0.0;
    H[14][0] = -sx * sz + cx * s_y * cz; H[14][1] = -cx * s_y * sz - sx * cz; H[14][2] = 0.0;
  }
  if (cfg.precision == NDT_PRECISION_FP32)
    hipLaunchKernelGGL(derivative_kernel<true>, dim3(blocks), dim3(kBlock), 0, st, g, cfg, d_pts, n, T, hess,
                       d_partials);
  else
    hipLaunchKernelGGL(derivative_kernel<false>, dim3(blocks), dim3(kBlock), 0, st, g, cfg, d_pts, n, T, hess,
                       d_partials);
  hipLaunchKernelGGL(reduce_partials, dim3(1, 1), dim3(64), 0, st, d_partials, blocks, kAccVals + 1, d_out);
  return hipGetLastError();
}

hipError_t launch_score(const GridView& g, double d1, double d2, int kind, const float* d_pts, uint32_t n,
                        const double* d_poses, int m, double* d_partials, int blocks_per_pose, double* d_out,
                        hipStream_t st) {
  if (kind == kScorePerPoint) {
    hipLaunchKernelGGL(score_kernel, dim3(blocks_per_pose, 1), dim3(kBlock), 0, st, g, d1, d2, kind, d_pts, n, d_poses,
                       d_partials, d_out);
    return hipGetLastError();
  }
  double* sums = d_partials + (size_t)m * blocks_per_pose * kPartialStride;
  hipLaunchKernelGGL(score_kernel, dim3(blocks_per_pose, m), dim3(kBlock), 0, st, g, d1, d2, kind, d_pts, n, d_poses,
                     d_partials, nullptr);
  hipLaunchKernelGGL(reduce_partials, dim3(1, m), dim3(64), 0, st, d_partials, blocks_per_pose, 2, sums);
  hipLaunchKernelGGL(finish_scores, dim3((m + 255) / 256), dim3(256), 0, st, sums, m, d_out);
  return hipGetLastError();
}

hipError_t launch_lookup(const GridView& g, int mode, const float* d_q, uint32_t n, uint32_t max_k, uint32_t* d_counts,
                         int32_t* d_idx, hipStream_t st) {
  hipLaunchKernelGGL(lookup_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, g, mode, d_q, n, max_k,
                     d_counts, d_idx);
  return hipGetLastError();
}

}  // namespace ndt
