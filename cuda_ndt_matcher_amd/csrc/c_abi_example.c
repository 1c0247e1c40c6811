/* A consumer of libndt_hip.so written against include/ndt_hip.h alone, the
 * way a foreign-function shim (the Rust crate's) binds it: plain C, no HIP
 * headers, no Python.
 *
 * It builds a small map (three orthogonal planes and a box, points on a
 * jittered 0.25 m lattice), takes a scan from the map points within 15 m of a
 * sensor at a known pose (expressed in the sensor frame), and checks:
 *  - ndt_hip_align from a guess 0.3 m / 2 degrees off returns the known
 *    pose within 5 cm / 0.5 degrees;
 *  - ndt_hip_align_batch over three starts agrees with single alignments
 *    (same status and iterations, poses within 1e-9);
 *  - ndt_hip_evaluate_nvtl at the result matches the result's NVTL;
 *  - an empty scan is refused with NDT_ERR_EMPTY_INPUT and a message.
 * Exit status 0 on success; prints one line per check. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ndt_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static uint64_t lcg = 0x9E3779B97F4A7C15ull;
static double urand(void) { /* uniform in [0, 1) */
  lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
}

/* R = Rz(yaw) Ry(pitch) Rx(roll), row-major. */
static void rot(const double p[6], double R[9]) {
  const double cr = cos(p[3]), sr = sin(p[3]), cp = cos(p[4]), sp = sin(p[4]), cy = cos(p[5]), sy = sin(p[5]);
  R[0] = cy * cp; R[1] = cy * sp * sr - sy * cr; R[2] = cy * sp * cr + sy * sr;
  R[3] = sy * cp; R[4] = sy * sp * sr + cy * cr; R[5] = sy * sp * cr - cy * sr;
  R[6] = -sp;     R[7] = cp * sr;                R[8] = cp * cr;
}

static size_t add_plane(float* out, size_t k, int axis, double c, double a0, double a1, double b0, double b1) {
  for (double a = a0; a < a1; a += 0.25)
    for (double b = b0; b < b1; b += 0.25) {
      double p[3];
      const double ja = a + 0.2 * (urand() - 0.5), jb = b + 0.2 * (urand() - 0.5), jc = c + 0.02 * (urand() - 0.5);
      if (axis == 2) { p[0] = ja; p[1] = jb; p[2] = jc; }
      else if (axis == 0) { p[0] = jc; p[1] = ja; p[2] = jb; }
      else { p[0] = ja; p[1] = jc; p[2] = jb; }
      out[3 * k] = (float)p[0];
      out[3 * k + 1] = (float)p[1];
      out[3 * k + 2] = (float)p[2];
      ++k;
    }
  return k;
}

static int fails = 0;
static void check(int ok, const char* what) {
  printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++fails;
}

int main(void) {
  const size_t cap = 200000;
  float* map = (float*)malloc(cap * 3 * sizeof(float));
  size_t n = 0;
  n = add_plane(map, n, 2, 0.0, -30, 30, -30, 30);    /* ground z = 0 */
  n = add_plane(map, n, 0, -12.0, -30, 30, 0, 8);     /* wall x = -12 */
  n = add_plane(map, n, 1, 9.0, -30, 30, 0, 6);       /* wall y = 9 */
  n = add_plane(map, n, 0, 4.0, -6, -2, 0, 3);        /* box faces */
  n = add_plane(map, n, 1, -2.0, 4, 8, 0, 3);
  n = add_plane(map, n, 2, 3.0, 4, 8, -6, -2);

  ndt_config cfg;
  ndt_hip_config_default(&cfg);
  ndt_handle* h = NULL;
  if (ndt_hip_create(&cfg, &h) != NDT_OK) {
    printf("FAIL create\n");
    return 1;
  }
  check(ndt_hip_set_target(h, map, n) == NDT_OK, "set_target");
  size_t voxels = 0;
  check(ndt_hip_num_voxels(h, &voxels) == NDT_OK && voxels > 100, "num_voxels");

  /* scan: map points within 15 m of the sensor, in the sensor frame */
  const double truth[6] = {1.5, -2.0, 1.2, 0.0, 0.0, 0.35};
  double R[9];
  rot(truth, R);
  float* scan = (float*)malloc(n * 3 * sizeof(float));
  size_t m = 0;
  for (size_t i = 0; i < n; i += 3) {
    const double d[3] = {map[3 * i] - truth[0], map[3 * i + 1] - truth[1], map[3 * i + 2] - truth[2]};
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > 15.0 * 15.0) continue;
    for (int r = 0; r < 3; ++r) scan[3 * m + r] = (float)(R[r] * d[0] + R[3 + r] * d[1] + R[6 + r] * d[2]);
    ++m;
  }
  printf("map %zu points, %zu voxels; scan %zu points\n", n, voxels, m);

  const double guess[6] = {1.8, -1.8, 1.2, 0.0, 0.0, 0.35 + 2.0 * M_PI / 180.0};
  ndt_result r;
  memset(&r, 0, sizeof(r));
  const int arc = ndt_hip_align(h, scan, m, guess, &r);
  if (arc != NDT_OK) printf("     ndt_hip_align returned %d: %s\n", arc, ndt_hip_last_error(h));
  check(arc == NDT_OK, "align");
  double et = 0.0, er = 0.0;
  for (int i = 0; i < 3; ++i) et = fmax(et, fabs(r.pose[i] - truth[i]));
  for (int i = 3; i < 6; ++i) er = fmax(er, fabs(r.pose[i] - truth[i]));
  printf("     status %d, %u iterations, error %.4f m %.4f deg, nvtl %.4f\n", r.status, r.iterations, et,
         er * 180.0 / M_PI, r.nvtl);
  check(et < 0.05 && er < 0.5 * M_PI / 180.0, "align recovers the sensor pose");

  double starts[18];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 6; ++i) starts[6 * k + i] = guess[i] + (i < 2 ? 0.1 * k : 0.0);
  ndt_result batch[3];
  check(ndt_hip_align_batch(h, scan, m, starts, 3, batch) == NDT_OK, "align_batch");
  int same = 1;
  for (int k = 0; k < 3; ++k) {
    ndt_result one;
    if (ndt_hip_align(h, scan, m, starts + 6 * k, &one) != NDT_OK) same = 0;
    if (one.status != batch[k].status || one.iterations != batch[k].iterations) same = 0;
    for (int i = 0; i < 6; ++i)
      if (fabs(one.pose[i] - batch[k].pose[i]) > 1e-9) same = 0;
  }
  check(same, "align_batch equals single alignments");

  double nv = 0.0;
  check(ndt_hip_evaluate_nvtl(h, scan, m, r.pose, &nv) == NDT_OK && fabs(nv - r.nvtl) < 1e-9,
        "evaluate_nvtl at the result equals the result's nvtl");

  const int rc = ndt_hip_align(h, scan, 0, guess, &r);
  check(rc == NDT_ERR_EMPTY_INPUT && strlen(ndt_hip_last_error(h)) > 0, "empty scan refused");

  ndt_hip_destroy(h);
  free(scan);
  free(map);
  printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
