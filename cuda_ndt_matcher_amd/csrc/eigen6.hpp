// Symmetric 6x6 eigendecomposition for the Newton step, written so that
// every lane of a wave runs the same scalar recurrence on its own registers
// and accumulates one row of the eigenvector matrix.  Nothing in the inner
// loops crosses lanes.
//
// Method: Householder reduction to tridiagonal form (four reflections), then
// implicit QL with the shift of EISPACK's tql2 (Bowdler, Martin, Reinsch &
// Wilkinson, Numer. Math. 11 (1968) 293-306), the small sub-diagonal test
// re-evaluated before every QL step.  Both steps are backward stable, so the
// eigenvalues agree with any other stable method to a few ulps of ||H|| and
// eigenvector k to that over its gap.
//
// The right-hand side b rides along: every reflection and rotation applied
// to the basis is applied to b as well, so y = Q^T b (Q: eigenvectors as
// columns) is available on every lane without a cross-lane reduction, and
// the row `row` of Q is the only per-lane part.  With z_k = y_k / lambda'_k
// the solution of the (clamped) system is component `row` = sum_k Q[row][k]
// z_k.
//
// The arithmetic policy M supplies
//   M::rsqrt(x), M::rcp(x): 1/sqrt(x) and 1/x for x > 0 to within a few ulps
//     (on the device: the hardware estimate plus one third-order step, much
//     shorter dependent chains than IEEE sqrt, hypot and division);
//   M::uni(v): a value every lane holds identically, as a wave-uniform one
//     (on the device: readfirstlane, so loop exits are scalar branches).
// Every rotation is c = p / r, s = e / r with r = sqrt(p^2 + e^2) formed from
// one reciprocal square root; no operand here comes near overflow (Hessian
// entries are far below 1e150).
#pragma once

#include <cmath>

namespace ndt {

// One third-order step each for 1/sqrt(x) and 1/x from an estimate y with a
// relative error up to about 2^-23 (the hardware's): with r = 1 - x y^2,
// 1/sqrt(x) = y (1 - r)^(-1/2) = y (1 + r/2 + 3r^2/8 + O(r^3)), and with
// e = 1 - x y, 1/x = y / (1 - e) = y (1 + e + e^2 + O(e^3)); the dropped terms
// are below 2^-66, so one step reaches a few ulps with a shorter dependent
// chain than two Newton steps.  x > 0 and finite.
__host__ __device__ inline double refine_rsqrt(double x, double y) {
  const double r = fma(-x * y, y, 1.0);
  const double t = r * fma(0.375, r, 0.5);
  return fma(y, t, y);
}
__host__ __device__ inline double refine_rcp(double x, double y) {
  const double e = fma(-x, y, 1.0);
  return fma(y, fma(e, e, e), y);
}

struct Eigen6Row {
  double lam[6];  // eigenvalues (unsorted)
  double y[6];    // Q^T b
  double q[6];    // row `row` of Q (zeros when row >= 6)
  int iters;      // QL steps taken
  int ok;         // 1; 0 when a QL block failed to converge within 30 steps;
                  // -1 when stop() ended the decomposition early
};

// stop() is polled before each QL block; when it returns true (the caller
// no longer needs the result) the decomposition ends with o.ok = -1.
template <class M, class Stop>
__host__ __device__ inline void eigen6_ql(const double* H, const double* b, int row, Eigen6Row& o, Stop stop) {
  // Full symmetric copy from the lower triangle.
  double a[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) a[i][j] = a[j][i] = H[i * 6 + j];
  double q[6], y[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    q[k] = (k == row) ? 1.0 : 0.0;
    y[k] = b[k];
  }

  // Householder: reflection k maps a[k+1..5][k] onto alpha e_{k+1}.
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double sig = 0.0;
#pragma unroll
    for (int j = k + 2; j < 6; ++j) sig += a[j][k] * a[j][k];
    if (M::uni(sig) == 0.0) continue;
    const double x1 = a[k + 1][k];
    const double nx = x1 * x1 + sig;
    const double mu = nx * M::rsqrt(nx);
    const double alpha = x1 <= 0.0 ? mu : -mu;  // no cancellation in x1 - alpha
    double v[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = j == k + 1 ? x1 - alpha : (j > k + 1 ? a[j][k] : 0.0);
    const double beta = 2.0 * M::rcp(v[k + 1] * v[k + 1] + sig);
    // A_sub <- P A_sub P with p = beta A_sub v, w = p - (beta/2)(v.p) v.
    double p[6], vp = 0.0;
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = k + 1; j < 6; ++j) s += a[i][j] * v[j];
      p[i] = beta * s;
      vp += v[i] * p[i];
    }
    const double K = 0.5 * beta * vp;
    double w[6];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) w[i] = p[i] - K * v[i];
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
#pragma unroll
      for (int j = k + 1; j <= i; ++j) {
        a[i][j] -= v[i] * w[j] + w[i] * v[j];
        a[j][i] = a[i][j];
      }
    a[k + 1][k] = a[k][k + 1] = alpha;
#pragma unroll
    for (int j = k + 2; j < 6; ++j) a[j][k] = a[k][j] = 0.0;
    // q <- q P (row of the accumulated basis), y <- P y.
    double qv = 0.0, yv = 0.0;
#pragma unroll
    for (int j = k + 1; j < 6; ++j) {
      qv += q[j] * v[j];
      yv += y[j] * v[j];
    }
    qv *= beta;
    yv *= beta;
#pragma unroll
    for (int j = k + 1; j < 6; ++j) {
      q[j] -= qv * v[j];
      y[j] -= yv * v[j];
    }
  }

  // Tridiagonal: d diagonal, e[i] couples i and i + 1, e[5] = 0.
  double d[6], e[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    d[i] = a[i][i];
    e[i] = i < 5 ? a[i + 1][i] : 0.0;
  }
  const double eps = 2.220446049250313e-16;
  double f = 0.0, tst1 = 0.0;
  int iters = 0, ok = 1;
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    if (M::uni(stop() ? 1 : 0)) {
      o.iters = iters;
      o.ok = -1;
      return;
    }
    tst1 = fmax(tst1, fabs(d[l]) + fabs(e[l]));
    for (int step = 0; step < 30; ++step) {
      // m: the first index >= l whose sub-diagonal is negligible (e[5] = 0).
      int m = 5;
#pragma unroll
      for (int mm = 4; mm >= l; --mm)
        if (fabs(e[mm]) <= eps * tst1) m = mm;
      m = M::uni(m);
      if (m == l) break;
      if (step == 29) ok = 0;
      ++iters;
      // Shift from the leading 2x2 of the block; the whole remaining
      // diagonal moves with it (f keeps the total).
      const double g0 = d[l];
      double p = (d[l + 1] - g0) * M::rcp(2.0 * e[l]);
      const double p1 = p * p + 1.0;
      double r = p1 * M::rsqrt(p1);
      if (p < 0.0) r = -r;
      d[l] = e[l] * M::rcp(p + r);
      d[l + 1] = e[l] * (p + r);
      const double dl1 = d[l + 1];
      const double h0 = g0 - d[l];
#pragma unroll
      for (int i = l + 2; i < 6; ++i) d[i] -= h0;
      f += h0;
      p = d[5];
#pragma unroll
      for (int mm = l + 1; mm < 5; ++mm)
        if (m == mm) p = d[mm];
      double c = 1.0, c2 = 1.0, c3 = 1.0, s = 0.0, s2 = 0.0;
      const double el1 = e[l + 1];
      // QL sweep from m - 1 down to l.
#pragma unroll
      for (int i = 4; i >= l; --i) {
        if (i >= m) continue;
        c3 = c2;
        c2 = c;
        s2 = s;
        const double g = c * e[i];
        const double h = c * p;
        const double x = p * p + e[i] * e[i];
        const double ir = M::rsqrt(x);
        e[i + 1] = s * (x * ir);
        s = e[i] * ir;
        c = p * ir;
        p = c * d[i] - s * g;
        d[i + 1] = h + s * (c * g + s * d[i]);
        double t = q[i + 1];
        q[i + 1] = s * q[i] + c * t;
        q[i] = c * q[i] - s * t;
        t = y[i + 1];
        y[i + 1] = s * y[i] + c * t;
        y[i] = c * y[i] - s * t;
      }
      p = -s * s2 * c3 * el1 * e[l] * M::rcp(dl1);
      e[l] = s * p;
      d[l] = c * p;
    }
    d[l] += f;
    e[l] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    o.lam[k] = d[k];
    o.y[k] = y[k];
    o.q[k] = q[k];
  }
  o.iters = iters;
  o.ok = ok;
}

}  // namespace ndt
