// Symmetric 3x3 eigendecomposition for the voxel covariances (cyclic Jacobi,
// one thread per matrix).  Rotations are formed with the arithmetic policy M
// of eigen6.hpp (M::rcp, M::rsqrt: 1/x and 1/sqrt(x) for x > 0 to within a
// few ulps; on the device the hardware estimate plus one third-order step,
// a much shorter sequence than IEEE division and square root).  A rotation
// whose angle parameter theta would square past 1e300 (an off-diagonal
// entry below 1e-150 of the diagonal gap) takes t = 1 / (2 theta) in IEEE
// arithmetic instead, its value to rounding.
#pragma once

#include <cmath>

namespace ndt {

// a_in: 3x3 row-major, symmetric (the lower triangle is read).  ev: the
// eigenvalues (unsorted); v: the eigenvectors as columns, row-major.
template <class M>
__host__ __device__ inline void sym_eigen3(const double* a_in, double* ev, double* v) {
  constexpr int N = 3;
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
        const double gap = a[q][q] - a[p][p];
        double t;
        if (fabs(gap) >= 1e150 * fabs(apq)) {
          t = apq / gap;  // theta = gap / (2 apq); t = 1 / (2 theta) to rounding
        } else {
          const double theta = gap * (0.5 * M::rcp(fabs(apq))) * (apq > 0.0 ? 1.0 : -1.0);
          const double x = fma(theta, theta, 1.0);
          const double root = x * M::rsqrt(x);  // sqrt(theta^2 + 1)
          t = (theta >= 0.0 ? 1.0 : -1.0) * M::rcp(fabs(theta) + root);
        }
        const double c = M::rsqrt(fma(t, t, 1.0));
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

}  // namespace ndt
