/* Host-code check of libndt_hip.so for the sanitizer build: a plain-C program
 * that drives every host-array entry point of include/ndt_hip.h, and their
 * error paths, so that AddressSanitizer and UndefinedBehaviorSanitizer see
 * the library's host code (argument checks, buffer growth, pinned staging,
 * result unpacking, the server's mailbox protocol, the MLP's parameter
 * copies).  Built twice by the Makefile: against the ordinary library and,
 * with the library's host code instrumented, as san/abi_host_check.
 *
 * The checks on results are light (status codes, sizes, agreement between
 * entry points that must agree); parity with the oracle is covered by the
 * Python tests.  Exit status 0 when every check passes. */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ndt_hip.h"

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define HOST_CHECK_ASAN 1
#endif
#endif

static uint64_t lcg = 0x2545F4914F6CDD1Dull;
static double urand(void) {
  lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
}

static int fails = 0;
static void check(int ok, const char* what) {
  printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  fflush(stdout);
  if (!ok) ++fails;
}

/* R = Rz(yaw) Ry(pitch) Rx(roll), row-major. */
static void rot(const double p[6], double R[9]) {
  const double cr = cos(p[3]), sr = sin(p[3]), cp = cos(p[4]), sp = sin(p[4]), cy = cos(p[5]), sy = sin(p[5]);
  R[0] = cy * cp; R[1] = cy * sp * sr - sy * cr; R[2] = cy * sp * cr + sy * sr;
  R[3] = sy * cp; R[4] = sy * sp * sr + cy * cr; R[5] = sy * sp * cr - cy * sr;
  R[6] = -sp;     R[7] = cp * sr;                R[8] = cp * cr;
}

/* Points on an axis-aligned rectangle at coordinate c of `axis`, on a
 * jittered 0.3 m lattice. */
static size_t plane(float* out, size_t k, int axis, double c, double a0, double a1, double b0, double b1) {
  for (double a = a0; a < a1; a += 0.3)
    for (double b = b0; b < b1; b += 0.3) {
      const double ja = a + 0.25 * (urand() - 0.5), jb = b + 0.25 * (urand() - 0.5), jc = c + 0.03 * (urand() - 0.5);
      float* p = out + 3 * k;
      if (axis == 2) { p[0] = (float)ja; p[1] = (float)jb; p[2] = (float)jc; }
      else if (axis == 0) { p[0] = (float)jc; p[1] = (float)ja; p[2] = (float)jb; }
      else { p[0] = (float)ja; p[1] = (float)jc; p[2] = (float)jb; }
      ++k;
    }
  return k;
}

/* Map points within `radius` of the sensor at `pose`, in the sensor frame;
 * every `stride`-th point. */
static size_t take_scan(const float* map, size_t n, const double pose[6], double radius, size_t stride, float* out) {
  double R[9];
  rot(pose, R);
  size_t m = 0;
  for (size_t i = 0; i < n; i += stride) {
    const double d[3] = {map[3 * i] - pose[0], map[3 * i + 1] - pose[1], map[3 * i + 2] - pose[2]};
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > radius * radius) continue;
    for (int r = 0; r < 3; ++r) out[3 * m + r] = (float)(R[r] * d[0] + R[3 + r] * d[1] + R[6 + r] * d[2]);
    ++m;
  }
  return m;
}

static double pose_err(const double a[6], const double b[6]) {
  double e = 0.0;
  for (int i = 0; i < 6; ++i) e = fmax(e, fabs(a[i] - b[i]));
  return e;
}

static int same_result(const ndt_result* a, const ndt_result* b, double tol) {
  return a->status == b->status && a->iterations == b->iterations && pose_err(a->pose, b->pose) <= tol;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int main(void) {
  const size_t cap = 160000;
  float* map = (float*)malloc(cap * 3 * sizeof(float));
  size_t n = 0;
  n = plane(map, n, 2, 0.0, -25, 25, -25, 25);
  n = plane(map, n, 0, -11.0, -25, 25, 0, 7);
  n = plane(map, n, 1, 8.0, -25, 25, 0, 5);
  n = plane(map, n, 0, 5.0, -7, -2, 0, 3);
  n = plane(map, n, 1, -3.0, 3, 8, 0, 3);
  n = plane(map, n, 2, 3.0, 3, 8, -7, -3);
  printf("map %zu points\n", n);
#ifdef HOST_CHECK_ASAN
  printf("built with AddressSanitizer\n");
#endif

  ndt_config cfg;
  ndt_hip_config_default(&cfg);
  ndt_handle* h = NULL;

  /* creation errors */
  check(ndt_hip_create(&cfg, NULL) == NDT_ERR_INVALID_ARGUMENT, "create without an out pointer refused");
  ndt_config bad = cfg;
  bad.resolution = 0.0f;
  check(ndt_hip_create(&bad, &h) != NDT_OK && h == NULL, "create with resolution 0 refused");
  bad = cfg;
  bad.search_mode = 9;
  check(ndt_hip_create(&bad, &h) != NDT_OK && h == NULL, "create with an unknown search mode refused");
  ndt_hip_destroy(NULL);

  if (ndt_hip_create(&cfg, &h) != NDT_OK) {
    printf("FAIL create\n");
    return 1;
  }
  const double truth[6] = {1.2, -1.5, 1.1, 0.01, -0.01, 0.3};
  float* scan = (float*)malloc(n * 3 * sizeof(float));
  const size_t m = take_scan(map, n, truth, 14.0, 3, scan);
  const double guess[6] = {1.4, -1.3, 1.1, 0.01, -0.01, 0.33};
  ndt_result r, r2;
  memset(&r, 0, sizeof(r));
  check(ndt_hip_has_target(h) == 0, "no target before set_target");
  check(ndt_hip_align(h, scan, m, guess, &r) == NDT_ERR_NO_TARGET && strlen(ndt_hip_last_error(h)) > 0,
        "align before set_target refused");
  check(ndt_hip_set_target(h, map, 0) == NDT_ERR_EMPTY_INPUT, "empty target refused");
  check(ndt_hip_set_target(h, NULL, 10) == NDT_ERR_INVALID_ARGUMENT, "null target refused");

  /* map builds: a subset first, then the whole map (device buffers grow) */
  check(ndt_hip_set_target(h, map, n / 2) == NDT_OK, "set_target (half map)");
  size_t v_half = 0, voxels = 0;
  check(ndt_hip_num_voxels(h, &v_half) == NDT_OK && v_half > 0, "num_voxels (half map)");
  check(ndt_hip_set_target(h, map, n) == NDT_OK, "set_target (whole map)");
  check(ndt_hip_num_voxels(h, &voxels) == NDT_OK && voxels > v_half && ndt_hip_has_target(h) == 1,
        "num_voxels grows with the map");
  printf("     %zu voxels, scan %zu points\n", voxels, m);
  {
    int32_t* coords = (int32_t*)malloc(voxels * 3 * sizeof(int32_t));
    float* means = (float*)malloc(voxels * 3 * sizeof(float));
    float* covs = (float*)malloc(voxels * 9 * sizeof(float));
    float* invs = (float*)malloc(voxels * 9 * sizeof(float));
    float* axes = (float*)malloc(voxels * 3 * sizeof(float));
    uint32_t* counts = (uint32_t*)malloc(voxels * sizeof(uint32_t));
    check(ndt_hip_get_voxels(h, coords, means, covs, invs, axes, counts) == NDT_OK, "get_voxels (all arrays)");
    check(ndt_hip_get_voxels(h, NULL, means, NULL, NULL, NULL, counts) == NDT_OK, "get_voxels (some arrays)");
    uint64_t total = 0;
    for (size_t i = 0; i < voxels; ++i) total += counts[i];
    check(total <= n && counts[0] >= cfg.min_points_per_voxel, "voxel counts");
    int32_t* keys = (int32_t*)malloc(n * 3 * sizeof(int32_t));
    uint32_t* order = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* skeys = (uint64_t*)malloc(n * sizeof(uint64_t));
    check(ndt_hip_get_build_debug(h, keys, order, skeys) == NDT_OK, "get_build_debug");
    int sorted = 1;
    for (size_t i = 1; i < n; ++i) sorted &= skeys[i - 1] <= skeys[i];
    check(sorted, "sorted keys ascend");
    uint32_t* lc = (uint32_t*)malloc(m * sizeof(uint32_t));
    int32_t* li = (int32_t*)malloc(m * 8 * sizeof(int32_t));
    for (int mode = 0; mode < 4; ++mode) {
      char what[64];
      snprintf(what, sizeof what, "lookup mode %d", mode);
      check(ndt_hip_lookup(h, mode, map, m, 8, lc, li) == NDT_OK, what);
    }
    check(ndt_hip_lookup(h, 7, map, m, 8, lc, li) == NDT_ERR_INVALID_ARGUMENT, "lookup with an unknown mode refused");
    free(lc); free(li); free(keys); free(order); free(skeys);
    free(coords); free(means); free(covs); free(invs); free(axes); free(counts);
  }

  /* derivatives, single and batched alignments, scoring */
  double s = 0.0, g[6], H[36], s2 = 0.0, g2[6];
  uint32_t corr = 0, corr2 = 0;
  check(ndt_hip_derivatives(h, scan, m, guess, 1, &s, g, H, &corr) == NDT_OK && corr > 0, "derivatives with Hessian");
  check(ndt_hip_derivatives(h, scan, m, guess, 0, &s2, g2, NULL, &corr2) == NDT_OK && corr2 == corr &&
        fabs(s2 - s) <= 1e-9 * fmax(1.0, fabs(s)), "derivatives without Hessian agree");
  check(ndt_hip_align(h, scan, m, guess, &r) == NDT_OK, "align");
  printf("     status %d, %u iterations, error %.4f, nvtl %.4f\n", r.status, r.iterations, pose_err(r.pose, truth),
         r.nvtl);
  check(pose_err(r.pose, truth) < 0.05, "align recovers the sensor pose");
  check(ndt_hip_align(h, scan, 0, guess, &r2) == NDT_ERR_EMPTY_INPUT, "empty scan refused");
  check(ndt_hip_align(h, NULL, m, guess, &r2) == NDT_ERR_EMPTY_INPUT, "null scan refused");
  check(ndt_hip_align(h, scan, m, guess, NULL) == NDT_ERR_INVALID_ARGUMENT, "null result refused");

  enum { kStarts = 8 };
  double starts[6 * kStarts];
  for (int k = 0; k < kStarts; ++k)
    for (int i = 0; i < 6; ++i) starts[6 * k + i] = guess[i] + (i < 2 ? 0.05 * (k - 4) : 0.0);
  ndt_result batch[kStarts];
  check(ndt_hip_align_batch(h, scan, m, starts, kStarts, batch) == NDT_OK, "align_batch");
  check(ndt_hip_align(h, scan, m, starts + 6 * 5, &r2) == NDT_OK && same_result(&r2, &batch[5], 1e-9),
        "align_batch equals a single alignment");
  check(ndt_hip_align_batch(h, scan, m, starts, 0, NULL) == NDT_OK, "align_batch of no poses is a no-op");

  {
    const double p1[6] = {-3.0, 2.0, 1.0, 0.0, 0.0, -0.4}, p2[6] = {6.0, 0.5, 1.3, 0.0, 0.0, 1.2};
    float* a = (float*)malloc(n * 3 * sizeof(float));
    float* b = (float*)malloc(n * 3 * sizeof(float));
    const size_t na = take_scan(map, n, p1, 12.0, 5, a), nb = take_scan(map, n, p2, 9.0, 17, b);
    const float* scans[3] = {scan, a, b};
    const size_t ns[3] = {m, na, nb};
    double ps[18];
    memcpy(ps, guess, sizeof guess);
    for (int i = 0; i < 6; ++i) ps[6 + i] = p1[i] + (i == 0 ? 0.1 : 0.0);
    for (int i = 0; i < 6; ++i) ps[12 + i] = p2[i] + (i == 1 ? -0.1 : 0.0);
    ndt_result out[3];
    check(ndt_hip_align_scans(h, scans, ns, ps, 3, out) == NDT_OK, "align_scans (uneven scans)");
    check(ndt_hip_align(h, a, na, ps + 6, &r2) == NDT_OK && same_result(&r2, &out[1], 1e-9),
          "align_scans equals a single alignment");
    printf("     scans of %zu, %zu, %zu points\n", m, na, nb);
    free(a);
    free(b);
  }

  double nv = 0.0, tp = 0.0;
  check(ndt_hip_evaluate_nvtl(h, scan, m, r.pose, &nv) == NDT_OK && fabs(nv - r.nvtl) < 1e-9, "evaluate_nvtl");
  check(ndt_hip_evaluate_transform_probability(h, scan, m, r.pose, &tp) == NDT_OK && tp > 0.0 && isfinite(tp),
        "evaluate_transform_probability");
  double* pp = (double*)malloc(m * sizeof(double));
  check(ndt_hip_per_point_scores(h, scan, m, r.pose, pp) == NDT_OK, "per_point_scores");
  free(pp);
  double nvb[kStarts];
  check(ndt_hip_evaluate_nvtl_batch(h, scan, m, starts, kStarts, nvb) == NDT_OK, "evaluate_nvtl_batch");

  double cxy[4], c6[36];
  check(ndt_hip_laplace_covariance_xy(r.hessian, cxy) == NDT_OK && cxy[0] > 0.0, "laplace_covariance_xy");
  check(ndt_hip_laplace_covariance_6x6(r.hessian, c6) == NDT_OK && c6[0] > 0.0, "laplace_covariance_6x6");

  /* sensor filter and downsampling */
  {
    ndt_filter_params fp;
    ndt_hip_filter_params_default(&fp);
    fp.min_distance = 1.0f;
    fp.max_distance = 20.0f;
    fp.min_z = -2.0f;
    fp.max_z = 6.0f;
    float* fo = (float*)malloc(m * 3 * sizeof(float));
    size_t fn = 0;
    ndt_filter_counts fc;
    check(ndt_hip_filter_points(h, scan, m, &fp, fo, m, &fn, &fc) == NDT_OK && fn == fc.kept &&
          fc.kept + fc.removed_by_distance + fc.removed_by_z == m, "filter_points counts add up");
    fp.downsample = 1;
    fp.downsample_resolution = 0.5f;
    size_t fd = 0;
    check(ndt_hip_filter_points(h, scan, m, &fp, fo, m, &fd, &fc) == NDT_OK && fd < fn, "filter_points downsamples");
    check(ndt_hip_filter_points(h, scan, m, &fp, fo, 1, &fd, &fc) == NDT_ERR_INVALID_ARGUMENT && fd > 1,
          "filter_points with a short output reports the needed size");
    size_t dn = 0;
    check(ndt_hip_voxel_downsample(h, scan, m, 0.4f, fo, m, &dn) == NDT_OK && dn > 0 && dn < m, "voxel_downsample");
    check(ndt_hip_voxel_downsample(h, scan, m, 0.0f, fo, m, &dn) == NDT_OK && dn == m,
          "voxel_downsample at resolution 0 keeps every point");
    free(fo);
  }

  /* regularisation pose on and off */
  ndt_hip_destroy(h);
  h = NULL;
  ndt_config rc = cfg;
  rc.regularization_enabled = 1;
  check(ndt_hip_create(&rc, &h) == NDT_OK && ndt_hip_set_target(h, map, n) == NDT_OK, "regularised handle");
  ndt_result plain, reg;
  check(ndt_hip_align(h, scan, m, guess, &plain) == NDT_OK, "align without a regularisation pose");
  const double rp[6] = {truth[0] + 0.3, truth[1] - 0.2, truth[2], 0, 0, truth[5]};
  check(ndt_hip_set_regularization_pose(h, rp) == NDT_OK && ndt_hip_align(h, scan, m, guess, &reg) == NDT_OK,
        "align with a regularisation pose");
  check(ndt_hip_clear_regularization_pose(h) == NDT_OK && ndt_hip_align(h, scan, m, guess, &r2) == NDT_OK &&
        same_result(&r2, &plain, 0.0), "clearing the pose restores the plain result");

  /* continuous-batching server */
  {
    ndt_server* srv = NULL;
    check(ndt_hip_server_start(h, 2, 4, m, &srv) == NDT_OK && srv != NULL, "server_start");
    if (srv) {
      ndt_server* second = NULL;
      check(ndt_hip_server_start(h, 1, 4, m, &second) != NDT_OK, "second server on the handle refused");
      check(ndt_hip_align(h, scan, m, guess, &r2) == NDT_ERR_INVALID_ARGUMENT, "align while serving refused");
      check(ndt_hip_server_set_idle_timeout(srv, 0.0) == NDT_ERR_INVALID_ARGUMENT &&
            ndt_hip_server_set_idle_timeout(srv, 30.0) == NDT_OK, "set_idle_timeout range");
      enum { kReq = 6 };
      uint64_t tickets[kReq];
      int ok = 1;
      for (int k = 0; k < kReq; ++k) {
        const size_t len = m / (size_t)(1 + k % 3);
        ok &= ndt_hip_server_submit(srv, scan, len, starts + 6 * k, &tickets[k]) == NDT_OK;
      }
      check(ok, "server_submit");
      uint64_t big = 0;
      check(ndt_hip_server_submit(srv, map, m + 1, guess, &big) == NDT_ERR_CAPACITY, "oversized request refused");
      ndt_result got[kReq];
      uint64_t gt[kReq];
      size_t have = 0;
      const double deadline = now_s() + 60.0;
      while (have < kReq && now_s() < deadline) {
        size_t c = 0;
        if (ndt_hip_server_poll(srv, got + have, gt + have, kReq - have, &c) != NDT_OK) break;
        have += c;
      }
      check(have == kReq, "server finishes every request");
      check(ndt_hip_server_stop(srv) == NDT_OK, "server_stop");
      int match = have == kReq;
      for (size_t i = 0; i < have && match; ++i) {
        int found = 0;
        for (int k = 0; k < kReq; ++k) found |= gt[i] == tickets[k];
        match &= found && got[i].iterations >= 1;
      }
      check(match, "tickets name the submitted requests");
    }
  }
  ndt_hip_destroy(h);
  h = NULL;

  /* line search, DIRECT7, FP32 terms, graph replay */
  ndt_config xc = cfg;
  xc.use_line_search = 1;
  xc.search_mode = NDT_SEARCH_DIRECT7;
  xc.precision = NDT_PRECISION_FP32;
  xc.use_graph = 1;
  check(ndt_hip_create(&xc, &h) == NDT_OK && ndt_hip_set_target(h, map, n) == NDT_OK, "line-search handle");
  ndt_result a1, a2;
  check(ndt_hip_align(h, scan, m, guess, &a1) == NDT_OK && ndt_hip_align(h, scan, m, guess, &a2) == NDT_OK &&
        same_result(&a1, &a2, 0.0), "graph replays equal the first launch");
  ndt_timing tm;
  check(ndt_hip_get_timing(h, &tm) == NDT_OK && tm.align_ms > 0.0f && tm.align_graph == 1, "get_timing");
  ndt_hip_destroy(h);
  h = NULL;

  /* phase stamps (direct launches) */
  ndt_config pc = cfg;
  pc.profile_phases = 1;
  ndt_phase_profile ph;
  check(ndt_hip_create(&pc, &h) == NDT_OK && ndt_hip_set_target(h, map, n) == NDT_OK &&
        ndt_hip_align(h, scan, m, guess, &a1) == NDT_OK, "phase-profiled alignment");
  check(ndt_hip_get_phase_profile(h, &ph) == NDT_OK && ph.iterations == a1.iterations && ph.total_us > 0.0f,
        "get_phase_profile");
  ndt_hip_destroy(h);
  h = NULL;

  /* autotuner: the first call of a bucket times candidate counts */
  ndt_config tc = cfg;
  tc.autotune = 1;
  check(ndt_hip_create(&tc, &h) == NDT_OK && ndt_hip_set_target(h, map, n) == NDT_OK, "autotuned handle");
  ndt_tuning_entry te[8];
  size_t tn = 0;
  check(ndt_hip_align(h, scan, m, guess, &a1) == NDT_OK && ndt_hip_align(h, scan, m, guess, &a2) == NDT_OK,
        "autotuned alignments");
  check(ndt_hip_get_tuning(h, te, 8, &tn) == NDT_OK && tn == 1 && te[0].workgroups_per_alignment >= 1, "get_tuning");
  check(ndt_hip_get_timing(h, &tm) == NDT_OK && tm.align_tuned == 1 &&
        tm.align_workgroups == (uint32_t)te[0].workgroups_per_alignment, "second call uses the tuned count");
  check(ndt_hip_get_tuning(h, te, 0, &tn) == NDT_OK && tn == 1, "get_tuning with no room");
  ndt_hip_destroy(h);
  h = NULL;

  /* proposal MLP: parameters round trip, forward/backward on host inputs */
  {
    const int32_t sizes[3] = {64, 32, 32};
    ndt_mlp* net = NULL;
    const int32_t odd[2] = {64, 33};
    check(ndt_hip_mlp_create(odd, 2, 1, &net) != NDT_OK && net == NULL, "mlp widths off 32 refused");
    check(ndt_hip_mlp_create(sizes, 3, 7, &net) == NDT_OK, "mlp_create");
    const size_t np = ndt_hip_mlp_num_params(net);
    check(np == (size_t)(64 * 32 + 32 + 32 * 32 + 32), "mlp_num_params");
    float* w = (float*)malloc(np * sizeof(float));
    float* w2 = (float*)malloc(np * sizeof(float));
    check(ndt_hip_mlp_get_params(net, w) == NDT_OK && ndt_hip_mlp_set_params(net, w) == NDT_OK &&
          ndt_hip_mlp_get_params(net, w2) == NDT_OK && memcmp(w, w2, np * sizeof(float)) == 0,
          "mlp parameters round trip");
    float* x = (float*)malloc(32 * 64 * sizeof(float));
    float* dout = (float*)malloc(32 * 32 * sizeof(float));
    float* out = (float*)malloc(32 * 32 * sizeof(float));
    for (int i = 0; i < 32 * 64; ++i) x[i] = (float)(urand() - 0.5);
    for (int i = 0; i < 32 * 32; ++i) dout[i] = (float)(urand() - 0.5);
    check(ndt_hip_mlp_forward_backward(net, x, 32, dout, out, w2) == NDT_OK && isfinite(out[0]) && isfinite(w2[0]),
          "mlp_forward_backward");
    check(ndt_hip_mlp_forward_backward(net, x, 31, dout, out, w2) == NDT_ERR_INVALID_ARGUMENT,
          "mlp batch off 32 refused");
    free(x); free(dout); free(out); free(w); free(w2);
    ndt_hip_mlp_destroy(net);
    ndt_hip_mlp_destroy(NULL);
  }

  free(scan);
  free(map);
  printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
