// Host build of the Newton solve's eigendecomposition (csrc/eigen6.hpp) and
// of the voxel covariances' 3x3 Jacobi (csrc/eigen3.hpp) for the CPU tests:
// the same sources the device compiles.  Built with -D__host__= -D__device__=
// by tests/test_eigen6.py and tests/test_eigen3.py.
#include "../../cuda_ndt_matcher_amd/csrc/eigen6.hpp"
#include "../../cuda_ndt_matcher_amd/csrc/eigen3.hpp"

#include <cstdint>
#include <cstring>

namespace {

// Exact host arithmetic.
struct HostMath {
  static double rsqrt(double x) { return 1.0 / std::sqrt(x); }
  static double rcp(double x) { return 1.0 / x; }
  template <class T>
  static T uni(T x) {
    return x;
  }
};

// The device's form: an estimate good to about 2^-22 (here the exact value
// with its low 30 mantissa bits cleared, a relative error up to 2^-22) and
// the third-order step of refine_rsqrt / refine_rcp.
double truncate22(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  b &= ~((uint64_t(1) << 30) - 1);
  std::memcpy(&v, &b, 8);
  return v;
}
struct EstimateMath {
  static double rsqrt(double x) { return ndt::refine_rsqrt(x, truncate22(1.0 / std::sqrt(x))); }
  static double rcp(double x) { return ndt::refine_rcp(x, truncate22(1.0 / x)); }
  template <class T>
  static T uni(T x) {
    return x;
  }
};

template <class M>
int full(const double* H, const double* b, double* lam, double* Q, double* y) {
  int iters = 0, ok = 1;
  for (int r = 0; r < 6; ++r) {
    ndt::Eigen6Row o;
    ndt::eigen6_ql<M>(H, b, r, o, [] { return false; });
    for (int k = 0; k < 6; ++k) Q[r * 6 + k] = o.q[k];
    if (r == 0)
      for (int k = 0; k < 6; ++k) {
        lam[k] = o.lam[k];
        y[k] = o.y[k];
      }
    iters = o.iters;
    ok &= o.ok;
  }
  return ok ? iters : -1;
}

}  // namespace

// mode 0: exact arithmetic; 1: the device's estimate-and-refine form.
extern "C" int eigen6_full(const double* H, const double* b, double* lam, double* Q, double* y, int mode) {
  return mode ? full<EstimateMath>(H, b, lam, Q, y) : full<HostMath>(H, b, lam, Q, y);
}

// 3x3: a row-major (lower triangle read), ev eigenvalues, v eigenvectors as
// columns; mode as above.
extern "C" void sym3_full(const double* a, double* ev, double* v, int mode) {
  if (mode)
    ndt::sym_eigen3<EstimateMath>(a, ev, v);
  else
    ndt::sym_eigen3<HostMath>(a, ev, v);
}

extern "C" double refine_rsqrt_from(double x, double y0) { return ndt::refine_rsqrt(x, y0); }
extern "C" double refine_rcp_from(double x, double y0) { return ndt::refine_rcp(x, y0); }
