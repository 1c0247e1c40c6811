// C ABI of libndt_hip.so (declared in include/ndt_hip.h).
//
// Host orchestration only: argument checks with the reference's error
// classes, grow-only device buffers owned by the handle, one stream per
// handle, and the launch shapes.  All arithmetic runs in the HIP kernels;
// there is no CPU fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ndt_hip.h"
#include "ndt_device.hpp"
#include "ndt_internal.hpp"

using ndt::DevBuf;

// Scratch of ndt_hip_align_device_noalloc: one alignment's job and exchange
// state (ndt::kNoAllocScratchBytes, about 131 KiB), held by every GPU the
// library runs on until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device array: the device address comes from
// hipGetSymbolAddress, once per device, into g_scratch below.
static __device__ __align__(256) unsigned char scratch[ndt::kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};

// Looks up this device's copy of `scratch` (device current).  Once per device.
hipError_t scratch_for_device(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}
}  // namespace

struct ndt_handle {
  ndt_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  ndt::DeviceGrid grid;
  ndt::DeviceGrid filt;  // scratch of the sensor point filter (never a map)
  DevBuf filt_in, filt_out;
  std::string err;
  DevBuf scan, jobs, results, ws, poses, partials, out, lookup_q, lookup_c, lookup_i, phase;
  ndt_phase_profile phase_profile{};
  int wall_clock_khz = 0;
  std::vector<const float*> scan_ptrs;
  ndt_result* host_results = nullptr;
  size_t host_results_cap = 0;
  bool has_reg_pose = false;
  double reg_t[3] = {0, 0, 0};
  hipEvent_t ev0{}, ev1{};
  ndt_timing timing{};
  int coop_capacity = 0;
  double d1 = 0, d2 = 0;
  std::vector<int32_t> slot_to_index;  // voxel slot -> exported index (-1 if dropped)
  ndt_server* server = nullptr;        // the running continuous-batching server, if any
  void* scratch = nullptr;             // this device's copy of the allocation-free scratch array
  ndt::SlotJob* host_jobs = nullptr;   // pinned staging for the jobs copy
  size_t host_jobs_cap = 0;
  // The last graph built for align* (use_graph) and what it was built for.
  struct GraphCache {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t m = 0;
    int W = 0;
    const void *jobs = nullptr, *results = nullptr, *ws = nullptr, *hjobs = nullptr, *hres = nullptr;
    ndt::NewtonConfig cfg;
    ndt::GridView view;
    void reset() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
  } gcache;
  // Autotuner: (bit length of the largest scan, bit length of the batch) ->
  // (workgroups per alignment, kernel ms).
  std::map<std::pair<int, int>, std::pair<int, float>> tuning;
};

namespace {

int set_err(ndt_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int hip_err(ndt_handle* h, hipError_t e, const char* what) {
  return set_err(h, NDT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NDT_HIP(call, what)                      \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return hip_err(h, e_, what); \
  } while (0)

// A cooperative launch (several workgroups per alignment) fails with
// hipErrorIllegalState while another process on the device has made
// cooperative launches: the runtime serves them to one process at a time, and
// a process keeps the resource until it exits.
int launch_err(ndt_handle* h, hipError_t e, const char* what, bool coop) {
  if (coop && e == hipErrorIllegalState)
    return set_err(h, NDT_ERR_HIP,
                   std::string(what) +
                       ": cooperative launch refused: another process on this GPU has made cooperative launches "
                       "(the runtime serves them to one process at a time, until that process exits); "
                       "workgroups_per_alignment = 1 needs no cooperative launch");
  return hip_err(h, e, what);
}

void gauss_params(double res, double outlier, double* d1, double* d2) {
  // GaussianParams::new (derivatives/types.rs:61-75)
  const double c1 = 10.0 * (1.0 - outlier);
  const double c2 = outlier / (res * res * res);
  const double d3 = -std::log(c2);
  *d1 = -std::log(c1 + c2) - d3;
  const double nom = -std::log(c1 * std::exp(-0.5) + c2) - d3;
  *d2 = -2.0 * std::log(nom / *d1);
}

ndt::NewtonConfig newton_config(const ndt_handle* h, int wg) {
  ndt::NewtonConfig c;
  std::memset(&c, 0, sizeof(c));  // padding included: the graph cache compares bytes
  c.d1 = h->d1;
  c.d2 = h->d2;
  c.trans_eps = h->cfg.trans_epsilon;
  c.step_size = h->cfg.step_size;
  c.reg_scale = h->cfg.regularization_scale_factor;
  for (int i = 0; i < 3; ++i) c.reg_t[i] = h->reg_t[i];
  c.max_iter = h->cfg.max_iterations;
  c.search_mode = h->cfg.search_mode;
  c.metric = h->cfg.distance_metric;
  c.use_ls = h->cfg.use_line_search;
  c.reg_active = (h->cfg.regularization_enabled && h->has_reg_pose) ? 1 : 0;
  c.precision = h->cfg.precision;
  c.wg_per_slot = wg;
  c.phase = h->cfg.profile_phases ? reinterpret_cast<unsigned long long*>(h->phase.ptr) : nullptr;
  return c;
}

// A running server holds every co-resident workgroup of the device, so any
// other launch from this handle would queue behind it while the caller still
// has to poll it: refused instead.
int check_no_server(ndt_handle* h) {
  if (h->server)
    return set_err(h, NDT_ERR_INVALID_ARGUMENT,
                   "a continuous-batching server is running on this handle; stop it first");
  return NDT_OK;
}

int check_ready(ndt_handle* h) {
  if (!h) return NDT_ERR_INVALID_ARGUMENT;
  if (int rc = check_no_server(h)) return rc;
  if (!h->grid.ready) return set_err(h, NDT_ERR_NO_TARGET, "No target set. Call set_target() first.");
  NDT_HIP(hipSetDevice(h->device), "hipSetDevice");
  return NDT_OK;
}

int upload_points(ndt_handle* h, DevBuf& buf, const float* xyz, size_t n, const float** dptr) {
  NDT_HIP(buf.reserve(n * 3 * sizeof(float)), "alloc points");
  NDT_HIP(hipMemcpyAsync(buf.ptr, xyz, n * 3 * sizeof(float), hipMemcpyHostToDevice, h->stream), "copy points");
  *dptr = reinterpret_cast<const float*>(buf.ptr);
  return NDT_OK;
}

int ensure_host_results(ndt_handle* h, size_t m) {
  if (m <= h->host_results_cap) return NDT_OK;
  if (h->host_results) (void)hipHostFree(h->host_results);
  h->host_results = nullptr;
  NDT_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->host_results), m * sizeof(ndt_result),
                        hipHostMallocCoherent | hipHostMallocMapped),  // written by the Newton kernel
          "pinned alloc");
  h->host_results_cap = m;
  return NDT_OK;
}

// Workgroups per alignment: about one point per thread, capped so that all
// slots' workgroups are co-resident (the slot barrier needs that).
int choose_wg(const ndt_handle* h, size_t n_max, size_t m) {
  const int cap = std::max(1, h->coop_capacity);
  int per_slot_cap = (int)std::max<size_t>(1, (size_t)cap / std::max<size_t>(1, m));
  if (h->cfg.workgroups_per_alignment > 0) return std::min(h->cfg.workgroups_per_alignment, per_slot_cap);
  // One point per thread: more workgroups shorten the lookups (each CU's
  // gather rate bounds them) but lengthen the slot exchange; a second point
  // per thread doubles the critical path.
  const size_t want = (n_max + 255) / 256;
  return (int)std::max<size_t>(1, std::min<size_t>(want, (size_t)per_slot_cap));
}

int bit_length(size_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// One launch of the Newton kernel over the m jobs staged in h->host_jobs,
// with W workgroups per alignment; results land in h->host_results and the
// kernel's device time in *ms.  With use_graph (and graph allowed), replays
// the cached graph, rebuilding it when the shape, a buffer or the
// configuration changed.
int launch_once(ndt_handle* h, size_t m, int W, bool allow_graph, float* ms, bool* used_graph) {
  // Several workgroups per alignment exchange partials through write-once
  // buffers, one per exchange the alignment can make.
  const int E = W > 1 ? ndt::newton_exchanges(h->cfg.max_iterations, h->cfg.use_line_search != 0) : 0;
  NDT_HIP(h->ws.reserve(ndt::newton_workspace_bytes((int)m, W, E)), "alloc workspace");
  if (h->cfg.profile_phases) NDT_HIP(h->phase.reserve(16 * sizeof(unsigned long long)), "alloc phase");
  ndt::NewtonConfig cfg = newton_config(h, W);  // after the reserves: it carries their pointers
  cfg.exchanges = E;
  const bool graph = allow_graph && h->cfg.use_graph && !h->cfg.profile_phases;
  *used_graph = graph;
  if (graph) {
    ndt_handle::GraphCache& c = h->gcache;
    const bool same = c.exec && c.m == m && c.W == W && c.jobs == h->jobs.ptr && c.results == h->results.ptr &&
                      c.ws == h->ws.ptr && c.hjobs == h->host_jobs && c.hres == h->host_results &&
                      std::memcmp(&c.cfg, &cfg, sizeof(cfg)) == 0 &&
                      std::memcmp(&c.view, &h->grid.view, sizeof(c.view)) == 0;
    if (!same) {
      c.reset();
      NDT_HIP(ndt::build_newton_graph(h->grid.view, cfg, reinterpret_cast<ndt::SlotJob*>(h->jobs.ptr), h->host_jobs,
                                      (int)m, reinterpret_cast<ndt_result*>(h->results.ptr), h->host_results,
                                      h->ws.ptr, h->ev0, h->ev1, &c.graph, &c.exec),
              "build graph");
      c.m = m;
      c.W = W;
      c.jobs = h->jobs.ptr;
      c.results = h->results.ptr;
      c.ws = h->ws.ptr;
      c.hjobs = h->host_jobs;
      c.hres = h->host_results;
      std::memcpy(&c.cfg, &cfg, sizeof(cfg));
      std::memcpy(&c.view, &h->grid.view, sizeof(c.view));
    }
    if (hipError_t e = hipGraphLaunch(c.exec, h->stream); e != hipSuccess)
      return launch_err(h, e, "graph launch", W > 1);
  } else {
    if (h->cfg.profile_phases)
      NDT_HIP(hipMemsetAsync(h->phase.ptr, 0, 16 * sizeof(unsigned long long), h->stream), "clear phase");
    // The kernel reads the jobs from and writes the results to the pinned,
    // mapped host buffers: no copies around the launch (each would add a
    // queue round trip to a launch of well under a millisecond).
    if (hipError_t e = ndt::launch_newton(h->grid.view, cfg, h->host_jobs, (int)m, h->host_results, h->ws.ptr,
                                          h->stream, W > 1, h->ev0);
        e != hipSuccess)
      return launch_err(h, e, "newton launch", W > 1);
    NDT_HIP(hipEventRecord(h->ev1, h->stream), "event");
  }
  NDT_HIP(hipStreamSynchronize(h->stream), "newton sync");
  NDT_HIP(hipEventElapsedTime(ms, h->ev0, h->ev1), "event time");
  return NDT_OK;
}

// Autotuner: times each candidate workgroup count on the caller's own jobs
// (one warm-up and two timed launches each, direct launches) and keeps the
// fastest for the (scan size, batch size) bucket.  Candidates: the default
// (about one point per thread) and the powers of two from an eighth of it to
// four times it, all within what keeps every slot's workgroups co-resident.
int autotune_wg(ndt_handle* h, size_t n_max, size_t m, int* W_out) {
  const int cap = std::max(1, h->coop_capacity);
  const int per_slot_cap = (int)std::max<size_t>(1, (size_t)cap / std::max<size_t>(1, m));
  const int h0 = (int)std::max<size_t>(1, std::min<size_t>((n_max + 255) / 256, (size_t)per_slot_cap));
  std::vector<int> cands = {h0};
  for (int w = 1; w <= std::min(per_slot_cap, 4 * h0); w *= 2)
    if (8 * w >= h0) cands.push_back(w);
  std::sort(cands.begin(), cands.end());
  cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
  int best_w = h0;
  float best_ms = 3.0e38f;
  for (int w : cands) {
    float ms = 0.f, t = 0.f;
    bool g = false;
    int rc = launch_once(h, m, w, false, &ms, &g);
    if (rc) return rc;
    float tmin = 3.0e38f;
    for (int r = 0; r < 2; ++r) {
      rc = launch_once(h, m, w, false, &t, &g);
      if (rc) return rc;
      tmin = std::min(tmin, t);
    }
    if (tmin < best_ms) {
      best_ms = tmin;
      best_w = w;
    }
  }
  h->tuning[{bit_length(n_max), bit_length(m)}] = {best_w, best_ms};
  *W_out = best_w;
  return NDT_OK;
}

// Runs m alignments: scan i = (d_pts[i], n[i]) from poses[6*i].
int run_alignments(ndt_handle* h, const float* const* d_pts, const size_t* n, const double* poses, size_t m,
                   ndt_result* out) {
  size_t n_max = 0;
  for (size_t i = 0; i < m; ++i) {
    if (n[i] == 0) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
    if (n[i] >= (1ull << 31)) return set_err(h, NDT_ERR_CAPACITY, "scan exceeds 2^31 points");
    n_max = std::max(n_max, n[i]);
  }
  int rc = ensure_host_results(h, m);
  if (rc) return rc;
  if (m > h->host_jobs_cap) {
    if (h->host_jobs) (void)hipHostFree(h->host_jobs);
    h->host_jobs = nullptr;
    h->host_jobs_cap = 0;
    NDT_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->host_jobs), m * sizeof(ndt::SlotJob),
                          hipHostMallocCoherent | hipHostMallocMapped),  // read by the Newton kernel
            "pinned alloc");
    h->host_jobs_cap = m;
  }
  for (size_t i = 0; i < m; ++i) {
    ndt::SlotJob& j = h->host_jobs[i];
    j.pts = d_pts[i];
    j.n = (uint32_t)n[i];
    j.pad = 0;
    for (int k = 0; k < 6; ++k) j.pose[k] = poses[6 * i + k];
  }
  NDT_HIP(h->jobs.reserve(m * sizeof(ndt::SlotJob)), "alloc jobs");
  NDT_HIP(h->results.reserve(m * sizeof(ndt_result)), "alloc results");
  int W = choose_wg(h, n_max, m);
  bool tuned = false;
  if (h->cfg.autotune && h->cfg.workgroups_per_alignment <= 0 && !h->cfg.profile_phases) {
    const auto it = h->tuning.find({bit_length(n_max), bit_length(m)});
    if (it != h->tuning.end()) {
      W = it->second.first;
    } else {
      rc = autotune_wg(h, n_max, m, &W);
      if (rc) return rc;
    }
    // a bucket spans sizes up to twice its smallest: keep the slots co-resident
    W = std::max(1, std::min(W, (int)std::max<size_t>(1, (size_t)std::max(1, h->coop_capacity) / m)));
    tuned = true;
  }
  if ((size_t)W * m > (size_t)std::max(1, h->coop_capacity) && W > 1)
    return set_err(h, NDT_ERR_CAPACITY, "too many cooperating workgroups for one launch");
  float ms = 0.f;
  bool graph = false;
  rc = launch_once(h, m, W, true, &ms, &graph);
  if (rc) return rc;
  h->timing.align_ms = ms;
  h->timing.align_workgroups = (uint32_t)(W * m);
  h->timing.align_slots = (uint32_t)m;
  h->timing.align_graph = graph ? 1u : 0u;
  h->timing.align_tuned = tuned ? 1u : 0u;
  std::memcpy(out, h->host_results, m * sizeof(ndt_result));
  if (h->cfg.profile_phases) {
    unsigned long long ph[16];
    NDT_HIP(hipMemcpy(ph, h->phase.ptr, sizeof(ph), hipMemcpyDeviceToHost), "copy phases");
    const double us = h->wall_clock_khz > 0 ? 1e3 / (double)h->wall_clock_khz : 0.0;
    ndt_phase_profile& p = h->phase_profile;
    p.derivative_us = (float)(ph[0] * us);
    p.exchange_us = (float)(ph[1] * us);
    p.solve_us = (float)(ph[2] * us);
    p.line_search_us = (float)(ph[3] * us);
    p.nvtl_us = (float)(ph[4] * us);
    p.total_us = (float)(ph[5] * us);
    p.iterations = (uint32_t)ph[6];
    p.line_search_evaluations = (uint32_t)ph[7];
    p.slowest_derivative_us = (float)(ph[8] * us);
    p.solve_setup_us = (float)(ph[9] * us);
    p.solve_step_us = (float)(ph[10] * us);
    p.eigen_solves = (uint32_t)ph[12];
    p.jacobi_sweeps = (uint32_t)ph[11];
  }
  return NDT_OK;
}

int run_score(ndt_handle* h, int kind, const float* d_pts, size_t n, const double* poses, size_t m, double* out) {
  NDT_HIP(h->poses.reserve(m * 6 * sizeof(double)), "alloc poses");
  NDT_HIP(hipMemcpyAsync(h->poses.ptr, poses, m * 6 * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy poses");
  const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 512));
  NDT_HIP(h->partials.reserve((m * blocks * 32 + 2 * m + 64) * sizeof(double)), "alloc partials");
  const size_t out_n = (kind == ndt::kScorePerPoint) ? n : m;
  NDT_HIP(h->out.reserve(out_n * sizeof(double)), "alloc out");
  NDT_HIP(hipEventRecord(h->ev0, h->stream), "event");
  NDT_HIP(ndt::launch_score(h->grid.view, h->d1, h->d2, kind, d_pts, (uint32_t)n,
                            reinterpret_cast<const double*>(h->poses.ptr), (int)m,
                            reinterpret_cast<double*>(h->partials.ptr), blocks,
                            reinterpret_cast<double*>(h->out.ptr), h->stream),
          "score launch");
  NDT_HIP(hipEventRecord(h->ev1, h->stream), "event");
  NDT_HIP(hipMemcpyAsync(out, h->out.ptr, out_n * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy scores");
  NDT_HIP(hipStreamSynchronize(h->stream), "score sync");
  float ms = 0.f;
  NDT_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1), "event time");
  h->timing.score_ms = ms;
  return NDT_OK;
}

int finish_target(ndt_handle* h, size_t n) {
  std::string err;
  const int rc = ndt::build_grid(h->grid, n, h->cfg.resolution, h->cfg.min_points_per_voxel, h->cfg.cell_tables != 0,
                                 h->stream, &err);
  h->timing.build_ms = h->grid.build_ms;
  h->timing.sort_ms = h->grid.sort_ms;
  if (rc != 0) {
    h->grid.ready = false;
    return set_err(h, rc, err);
  }
  // slot -> exported index map
  std::vector<uint32_t> valid(h->grid.segments);
  NDT_HIP(hipMemcpy(valid.data(), h->grid.valid.ptr, valid.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
          "copy valid flags");
  h->slot_to_index.assign(h->grid.segments, -1);
  int32_t k = 0;
  for (uint32_t s = 0; s < h->grid.segments; ++s)
    if (valid[s]) h->slot_to_index[s] = k++;
  h->grid.ready = true;
  return NDT_OK;
}

bool valid_config(const ndt_config* c, std::string* why) {
  if (!(c->resolution > 0.0f) || !std::isfinite(c->resolution)) return *why = "resolution must be positive", false;
  if (c->max_iterations > 256) return *why = "max_iterations above 256 is not supported", false;
  if (c->search_mode < 0 || c->search_mode > 3) return *why = "unknown search_mode", false;
  if (c->distance_metric < 0 || c->distance_metric > 1) return *why = "unknown distance_metric", false;
  if (c->precision < 0 || c->precision > 1) return *why = "unknown precision", false;
  if (!(c->outlier_ratio > 0.0 && c->outlier_ratio < 1.0)) return *why = "outlier_ratio must be in (0,1)", false;
  return true;
}

}  // namespace

extern "C" {

void ndt_hip_config_default(ndt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->resolution = 2.0f;
  c->max_iterations = 30;
  c->trans_epsilon = 0.01;
  c->step_size = 0.1;
  c->outlier_ratio = 0.55;
  c->regularization = 0.001;
  c->use_line_search = 0;
  c->min_points_per_voxel = 6;
  c->regularization_enabled = 0;
  c->regularization_scale_factor = 0.01;
  c->distance_metric = NDT_METRIC_POINT_TO_DISTRIBUTION;
  c->search_mode = NDT_SEARCH_KDTREE;
  c->precision = NDT_PRECISION_FP64;
  c->workgroups_per_alignment = 0;
  c->profile_phases = 0;
  c->cell_tables = 1;
  c->use_graph = 0;
  c->autotune = 0;
}

int ndt_hip_create(const ndt_config* cfg, ndt_handle** out) {
  if (!cfg || !out) return NDT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::string why;
  if (!valid_config(cfg, &why)) return NDT_ERR_INVALID_ARGUMENT;
  if (cfg->precision != NDT_PRECISION_FP64 && cfg->precision != NDT_PRECISION_FP32) return NDT_ERR_UNSUPPORTED;
  auto* h = new ndt_handle();
  h->cfg = *cfg;
  gauss_params((double)cfg->resolution, cfg->outlier_ratio, &h->d1, &h->d2);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipEventCreate(&h->grid.ev_build0);
  if (e == hipSuccess) e = hipEventCreate(&h->grid.ev_build1);
  if (e == hipSuccess) e = hipEventCreate(&h->grid.ev_sort0);
  if (e == hipSuccess) e = hipEventCreate(&h->grid.ev_sort1);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&h->grid.host_scalars), 64 * sizeof(uint32_t),
                      hipHostMallocCoherent | hipHostMallocMapped);  // written by kernels
  if (e != hipSuccess) {
    ndt_hip_destroy(h);
    return NDT_ERR_HIP;
  }
  h->coop_capacity = ndt::newton_max_coresident_workgroups();
  (void)hipDeviceGetAttribute(&h->wall_clock_khz, hipDeviceAttributeWallClockRate, h->device);
  if (scratch_for_device(h->device, &h->scratch) != hipSuccess) {
    ndt_hip_destroy(h);
    return NDT_ERR_HIP;
  }
  *out = h;
  return NDT_OK;
}

void ndt_hip_destroy(ndt_handle* h) {
  if (!h) return;
  if (h->server) (void)ndt_hip_server_stop(h->server);
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  ndt::DeviceGrid& G = h->grid;
  for (DevBuf* b : {&G.owned_points, &G.sort_keys_a, &G.sort_keys_b, &G.sort_vals_a, &G.sort_vals_b, &G.hist,
                    &G.tile_hist, &G.tile_heads, &G.seg_start, &G.scalars, &G.vox, &G.cov, &G.axis, &G.count, &G.coord,
                    &G.valid, &G.table, &G.bitmap, &G.occ, &G.cell_nbr, &G.cell_ent, &h->scan, &h->jobs, &h->results, &h->ws, &h->poses, &h->partials, &h->out,
                    &h->lookup_q, &h->lookup_c, &h->lookup_i, &h->phase})
    b->release();
  h->gcache.reset();
  ndt::DeviceGrid& F = h->filt;
  for (DevBuf* b : {&F.owned_points, &F.sort_keys_a, &F.sort_keys_b, &F.sort_vals_a, &F.sort_vals_b, &F.hist,
                    &F.tile_hist, &F.tile_heads, &F.seg_start, &F.scalars, &h->filt_in, &h->filt_out})
    b->release();
  for (ndt::DeviceGrid* g : {&G, &F}) {
    if (g->host_scalars) (void)hipHostFree(g->host_scalars);
    if (g->host_part) (void)hipHostFree(g->host_part);
    if (g->host_valid) (void)hipHostFree(g->host_valid);
  }
  for (hipEvent_t ev : {F.ev_sort0, F.ev_sort1})
    if (ev) (void)hipEventDestroy(ev);
  if (h->host_results) (void)hipHostFree(h->host_results);
  if (h->host_jobs) (void)hipHostFree(h->host_jobs);
  for (hipEvent_t ev : {h->ev0, h->ev1, G.ev_build0, G.ev_build1, G.ev_sort0, G.ev_sort1})
    if (ev) (void)hipEventDestroy(ev);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char* ndt_hip_last_error(const ndt_handle* h) { return h ? h->err.c_str() : "null handle"; }

int ndt_hip_set_target(ndt_handle* h, const float* xyz, size_t n) {
  if (!h) return NDT_ERR_INVALID_ARGUMENT;
  if (int rc = check_no_server(h)) return rc;
  if (n == 0) return set_err(h, NDT_ERR_EMPTY_INPUT, "Target point cloud is empty");
  if (!xyz) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null points");
  NDT_HIP(hipSetDevice(h->device), "hipSetDevice");
  const float* d = nullptr;
  int rc = upload_points(h, h->grid.owned_points, xyz, n, &d);
  if (rc) return rc;
  h->grid.points = d;
  return finish_target(h, n);
}

int ndt_hip_set_target_device(ndt_handle* h, const float* d_xyz, size_t n) {
  if (!h) return NDT_ERR_INVALID_ARGUMENT;
  if (int rc = check_no_server(h)) return rc;
  if (n == 0) return set_err(h, NDT_ERR_EMPTY_INPUT, "Target point cloud is empty");
  if (!d_xyz) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null points");
  NDT_HIP(hipSetDevice(h->device), "hipSetDevice");
  h->grid.points = d_xyz;
  return finish_target(h, n);
}

int ndt_hip_has_target(const ndt_handle* h) { return (h && h->grid.ready) ? 1 : 0; }

int ndt_hip_num_voxels(const ndt_handle* h, size_t* out) {
  if (!h || !out) return NDT_ERR_INVALID_ARGUMENT;
  *out = h->grid.ready ? h->grid.num_valid : 0;
  return NDT_OK;
}

int ndt_hip_get_voxels(ndt_handle* h, int32_t* coords, float* means, float* covs, float* invs, float* axes,
                       uint32_t* counts) {
  int rc = check_ready(h);
  if (rc) return rc;
  const uint32_t S = h->grid.segments;
  std::vector<float> vox(12 * (size_t)S), cov(9 * (size_t)S), ax(3 * (size_t)S);
  std::vector<uint32_t> cnt(S);
  std::vector<int32_t> crd(3 * (size_t)S);
  NDT_HIP(hipStreamSynchronize(h->stream), "sync");
  NDT_HIP(hipMemcpy(vox.data(), h->grid.vox.ptr, vox.size() * 4, hipMemcpyDeviceToHost), "copy voxels");
  NDT_HIP(hipMemcpy(cov.data(), h->grid.cov.ptr, cov.size() * 4, hipMemcpyDeviceToHost), "copy voxels");
  NDT_HIP(hipMemcpy(ax.data(), h->grid.axis.ptr, ax.size() * 4, hipMemcpyDeviceToHost), "copy voxels");
  NDT_HIP(hipMemcpy(cnt.data(), h->grid.count.ptr, cnt.size() * 4, hipMemcpyDeviceToHost), "copy voxels");
  NDT_HIP(hipMemcpy(crd.data(), h->grid.coord.ptr, crd.size() * 4, hipMemcpyDeviceToHost), "copy voxels");
  for (uint32_t s = 0; s < S; ++s) {
    const int32_t k = h->slot_to_index[s];
    if (k < 0) continue;
    const float* v = &vox[12 * (size_t)s];
    if (coords) std::memcpy(coords + 3 * (size_t)k, &crd[3 * (size_t)s], 12);
    if (means) std::memcpy(means + 3 * (size_t)k, v, 12);
    if (invs) std::memcpy(invs + 9 * (size_t)k, v + 3, 36);
    if (covs) std::memcpy(covs + 9 * (size_t)k, &cov[9 * (size_t)s], 36);
    if (axes) std::memcpy(axes + 3 * (size_t)k, &ax[3 * (size_t)s], 12);
    if (counts) counts[k] = cnt[s];
  }
  return NDT_OK;
}

int ndt_hip_get_build_debug(ndt_handle* h, int32_t* point_keys, uint32_t* sorted_index, uint64_t* sorted_keys) {
  int rc = check_ready(h);
  if (rc) return rc;
  const size_t n = h->grid.n_points;
  std::vector<uint64_t> keys(n);
  std::vector<uint32_t> idx(n);
  NDT_HIP(hipStreamSynchronize(h->stream), "sync");
  if (h->grid.key64) {
    NDT_HIP(hipMemcpy(keys.data(), h->grid.sorted_keys, n * 8, hipMemcpyDeviceToHost), "copy keys");
  } else {
    std::vector<uint32_t> k32(n);
    NDT_HIP(hipMemcpy(k32.data(), h->grid.sorted_keys, n * 4, hipMemcpyDeviceToHost), "copy keys");
    for (size_t i = 0; i < n; ++i) keys[i] = k32[i];
  }
  NDT_HIP(hipMemcpy(idx.data(), h->grid.sorted_vals, n * 4, hipMemcpyDeviceToHost), "copy order");
  if (sorted_index) std::memcpy(sorted_index, idx.data(), n * 4);
  if (sorted_keys) std::memcpy(sorted_keys, keys.data(), n * 8);
  if (point_keys) {
    const int* bits = h->grid.key_bits_axis;
    const int nb = std::max(bits[0], std::max(bits[1], bits[2]));
    for (size_t j = 0; j < n; ++j) {
      uint32_t r[3] = {0, 0, 0};
      int pos = 0;
      for (int i = 0; i < nb; ++i)
        for (int a = 0; a < 3; ++a)
          if (i < bits[a]) r[a] |= (uint32_t)((keys[j] >> pos++) & 1ull) << i;
      const size_t p = idx[j];
      for (int a = 0; a < 3; ++a) point_keys[3 * p + a] = (int32_t)((int64_t)r[a] + h->grid.key_min[a]);
    }
  }
  return NDT_OK;
}

int ndt_hip_lookup(ndt_handle* h, int32_t mode, const float* q, size_t n, uint32_t max_k, uint32_t* counts,
                   int32_t* indices) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (mode < 0 || mode > 4 || !q || !counts || (max_k && !indices))
    return set_err(h, NDT_ERR_INVALID_ARGUMENT, "bad lookup arguments");
  if (n == 0) return NDT_OK;
  const float* dq = nullptr;
  rc = upload_points(h, h->lookup_q, q, n, &dq);
  if (rc) return rc;
  NDT_HIP(h->lookup_c.reserve(n * 4), "alloc");
  NDT_HIP(h->lookup_i.reserve(n * std::max<uint32_t>(1, max_k) * 4), "alloc");
  NDT_HIP(ndt::launch_lookup(h->grid.view, mode, dq, (uint32_t)n, max_k, reinterpret_cast<uint32_t*>(h->lookup_c.ptr),
                             reinterpret_cast<int32_t*>(h->lookup_i.ptr), h->stream),
          "lookup launch");
  NDT_HIP(hipMemcpyAsync(counts, h->lookup_c.ptr, n * 4, hipMemcpyDeviceToHost, h->stream), "copy");
  if (max_k)
    NDT_HIP(hipMemcpyAsync(indices, h->lookup_i.ptr, n * max_k * 4, hipMemcpyDeviceToHost, h->stream), "copy");
  NDT_HIP(hipStreamSynchronize(h->stream), "lookup sync");
  for (size_t i = 0; i < n * max_k; ++i)
    if (indices[i] >= 0) indices[i] = h->slot_to_index[indices[i]];
  return NDT_OK;
}

int ndt_hip_derivatives(ndt_handle* h, const float* xyz, size_t n, const double pose[6], int32_t compute_hessian,
                        double* score, double grad[6], double hess[36], uint32_t* num_correspondences) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!xyz || !pose || n == 0) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1024));
  NDT_HIP(h->partials.reserve((size_t)blocks * 32 * sizeof(double)), "alloc");
  NDT_HIP(h->out.reserve(64 * sizeof(double)), "alloc");
  NDT_HIP(ndt::launch_derivatives(h->grid.view, newton_config(h, 1), d, (uint32_t)n, pose, compute_hessian,
                                  reinterpret_cast<double*>(h->partials.ptr), blocks,
                                  reinterpret_cast<double*>(h->out.ptr), h->stream),
          "derivative launch");
  double tot[32];
  NDT_HIP(hipMemcpyAsync(tot, h->out.ptr, 29 * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
  NDT_HIP(hipStreamSynchronize(h->stream), "sync");
  if (score) *score = tot[0];
  if (grad)
    for (int i = 0; i < 6; ++i) grad[i] = tot[1 + i];
  if (hess)
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j <= i; ++j) hess[i * 6 + j] = hess[j * 6 + i] = tot[7 + i * (i + 1) / 2 + j];
  if (num_correspondences) *num_correspondences = (uint32_t)tot[28];
  return NDT_OK;
}

int ndt_hip_align(ndt_handle* h, const float* xyz, size_t n, const double pose[6], ndt_result* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!pose || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0 || !xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  return run_alignments(h, &d, &n, pose, 1, out);
}

int ndt_hip_align_device(ndt_handle* h, const float* d_xyz, size_t n, const double pose[6], ndt_result* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!pose || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0 || !d_xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  return run_alignments(h, &d_xyz, &n, pose, 1, out);
}

int ndt_hip_align_device_noalloc(ndt_handle* h, const float* d_xyz, size_t n, const double pose[6],
                                 ndt_result* d_out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!pose || !d_out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0 || !d_xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (n >= (1ull << 31)) return set_err(h, NDT_ERR_CAPACITY, "scan exceeds 2^31 points");
  // Same workgroup count as ndt_hip_align_device (configured or about one
  // point per thread), without the autotuner, which launches extra trials.
  const int W = choose_wg(h, n, 1);
  if (W > ndt::kNoAllocMaxWorkgroups)
    return set_err(h, NDT_ERR_CAPACITY, "allocation-free alignment supports at most 256 workgroups per alignment");
  if (!h->scratch) return set_err(h, NDT_ERR_HIP, "allocation-free scratch is not set up on this device");
  ndt::SlotJob job;
  job.pts = d_xyz;
  job.n = (uint32_t)n;
  job.pad = 0;
  for (int k = 0; k < 6; ++k) job.pose[k] = pose[k];
  ndt::NewtonConfig cfg = newton_config(h, W);
  cfg.phase = nullptr;  // phase stamps need a buffer of their own
  auto* d_job = reinterpret_cast<ndt::SlotJob*>(h->scratch);
  void* ws = reinterpret_cast<char*>(h->scratch) + ndt::kNoAllocWorkspaceOffset;
  NDT_HIP(ndt::store_job(job, d_job, h->stream), "job store");
  // launch_newton clears the arrival counters with a kernel before the
  // Newton kernel; the partials are written before they are read.
  if (hipError_t e = ndt::launch_newton(h->grid.view, cfg, d_job, 1, d_out, ws, h->stream, W > 1, h->ev0);
      e != hipSuccess)
    return launch_err(h, e, "newton launch", W > 1);
  NDT_HIP(hipEventRecord(h->ev1, h->stream), "event");
  NDT_HIP(hipStreamSynchronize(h->stream), "newton sync");
  float ms = 0.f;
  NDT_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1), "event time");
  h->timing.align_ms = ms;
  h->timing.align_workgroups = (uint32_t)W;
  h->timing.align_slots = 1;
  h->timing.align_graph = 0;
  h->timing.align_tuned = 0;
  return NDT_OK;
}

int ndt_hip_align_batch(ndt_handle* h, const float* xyz, size_t n, const double* poses, size_t m, ndt_result* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0 || !xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (m == 0) return NDT_OK;
  if (!poses || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  std::vector<const float*> ptrs(m, d);
  std::vector<size_t> ns(m, n);
  return run_alignments(h, ptrs.data(), ns.data(), poses, m, out);
}

int ndt_hip_align_batch_device(ndt_handle* h, const float* d_xyz, size_t n, const double* poses, size_t m,
                               ndt_result* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0 || !d_xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (m == 0) return NDT_OK;
  if (!poses || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<const float*> ptrs(m, d_xyz);
  std::vector<size_t> ns(m, n);
  return run_alignments(h, ptrs.data(), ns.data(), poses, m, out);
}

int ndt_hip_align_scans(ndt_handle* h, const float* const* xyz, const size_t* n, const double* poses, size_t m,
                        ndt_result* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (m == 0) return NDT_OK;
  if (!xyz || !n || !poses || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  size_t total = 0;
  for (size_t i = 0; i < m; ++i) {
    if (n[i] == 0 || !xyz[i]) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
    total += n[i];
  }
  NDT_HIP(h->scan.reserve(total * 3 * sizeof(float)), "alloc scans");
  std::vector<const float*> ptrs(m);
  size_t off = 0;
  for (size_t i = 0; i < m; ++i) {
    float* dst = reinterpret_cast<float*>(h->scan.ptr) + 3 * off;
    NDT_HIP(hipMemcpyAsync(dst, xyz[i], n[i] * 3 * sizeof(float), hipMemcpyHostToDevice, h->stream), "copy scan");
    ptrs[i] = dst;
    off += n[i];
  }
  return run_alignments(h, ptrs.data(), n, poses, m, out);
}

int ndt_hip_evaluate_nvtl(ndt_handle* h, const float* xyz, size_t n, const double pose[6], double* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0 || !xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (!pose || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  return run_score(h, ndt::kScoreNvtl, d, n, pose, 1, out);
}

int ndt_hip_evaluate_transform_probability(ndt_handle* h, const float* xyz, size_t n, const double pose[6],
                                           double* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0 || !xyz) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (!pose || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  return run_score(h, ndt::kScoreTp, d, n, pose, 1, out);
}

int ndt_hip_per_point_scores(ndt_handle* h, const float* xyz, size_t n, const double pose[6], double* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0) return NDT_OK;  // ndt.rs:811-813 returns an empty vector
  if (!xyz || !pose || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  return run_score(h, ndt::kScorePerPoint, d, n, pose, 1, out);
}

int ndt_hip_evaluate_nvtl_batch(ndt_handle* h, const float* xyz, size_t n, const double* poses, size_t m,
                                double* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (m == 0) return NDT_OK;
  if (!poses || !out) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) {  // ndt.rs:846-848: zeros for an empty cloud
    for (size_t i = 0; i < m; ++i) out[i] = 0.0;
    return NDT_OK;
  }
  if (!xyz) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null points");
  const float* d = nullptr;
  rc = upload_points(h, h->scan, xyz, n, &d);
  if (rc) return rc;
  return run_score(h, ndt::kScoreNvtlCell27, d, n, poses, m, out);
}

void ndt_hip_filter_params_default(ndt_filter_params* p) {
  if (!p) return;
  p->min_distance = 0.0f;  // filtering/mod.rs:22-31
  p->max_distance = FLT_MAX;
  p->min_z = -FLT_MAX;
  p->max_z = FLT_MAX;
  p->downsample_resolution = 0.0f;
  p->downsample = 0;
}

namespace {

// filter_points on the handle's filter scratch: d_in (n points) -> d_out.
int run_filter(ndt_handle* h, const float* d_in, size_t n, const ndt_filter_params* p, float* d_out, size_t out_cap,
               size_t* out_n, ndt_filter_counts* counts) {
  ndt::DeviceGrid& F = h->filt;
  if (!F.ev_sort0) NDT_HIP(hipEventCreate(&F.ev_sort0), "event");
  if (!F.ev_sort1) NDT_HIP(hipEventCreate(&F.ev_sort1), "event");
  F.points = d_in;
  uint64_t c[4];
  std::string err;
  const int rc = ndt::filter_points(F, n, p->min_distance, p->max_distance, p->min_z, p->max_z, p->downsample,
                                    p->downsample_resolution, d_out, out_cap, out_n, c, h->stream, &err);
  if (counts) {
    counts->kept = c[0];
    counts->removed_by_distance = c[1];
    counts->removed_by_z = c[2];
    counts->removed_by_downsampling = c[3];
  }
  if (rc != 0) return set_err(h, rc, err);
  return NDT_OK;
}

int check_filter_args(ndt_handle* h, const void* in, size_t n, const ndt_filter_params* p, const void* out,
                      const size_t* out_n) {
  if (!h) return NDT_ERR_INVALID_ARGUMENT;
  if (int rc = check_no_server(h)) return rc;
  if (!p || !out_n || (n > 0 && (!in || !out))) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "null argument");
  NDT_HIP(hipSetDevice(h->device), "hipSetDevice");
  return NDT_OK;
}

}  // namespace

int ndt_hip_filter_points(ndt_handle* h, const float* xyz, size_t n, const ndt_filter_params* p, float* out,
                          size_t out_cap, size_t* out_n, ndt_filter_counts* counts) {
  if (int rc = check_filter_args(h, xyz, n, p, out, out_n)) return rc;
  if (n == 0) return run_filter(h, nullptr, 0, p, nullptr, 0, out_n, counts);
  const float* d = nullptr;
  if (int rc = upload_points(h, h->filt_in, xyz, n, &d)) return rc;
  NDT_HIP(h->filt_out.reserve(n * 3 * sizeof(float)), "alloc points");
  float* d_out = reinterpret_cast<float*>(h->filt_out.ptr);
  if (int rc = run_filter(h, d, n, p, d_out, n, out_n, counts)) return rc;
  if (out_cap < *out_n) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "output capacity below the result count");
  if (*out_n) {
    NDT_HIP(hipMemcpyAsync(out, d_out, *out_n * 3 * sizeof(float), hipMemcpyDeviceToHost, h->stream), "copy out");
    NDT_HIP(hipStreamSynchronize(h->stream), "filter sync");
  }
  return NDT_OK;
}

int ndt_hip_filter_points_device(ndt_handle* h, const float* d_xyz, size_t n, const ndt_filter_params* p,
                                 float* d_out, size_t out_cap, size_t* out_n, ndt_filter_counts* counts) {
  if (int rc = check_filter_args(h, d_xyz, n, p, d_out, out_n)) return rc;
  return run_filter(h, d_xyz, n, p, d_out, out_cap, out_n, counts);
}

int ndt_hip_voxel_downsample(ndt_handle* h, const float* xyz, size_t n, float resolution, float* out, size_t out_cap,
                             size_t* out_n) {
  // voxel_downsample_cpu alone: no distance or z test removes anything.
  ndt_filter_params p;
  p.min_distance = 0.0f;
  p.max_distance = INFINITY;
  p.min_z = -INFINITY;
  p.max_z = INFINITY;
  p.downsample_resolution = resolution;
  p.downsample = 1;
  return ndt_hip_filter_points(h, xyz, n, &p, out, out_cap, out_n, nullptr);
}

int ndt_hip_set_regularization_pose(ndt_handle* h, const double pose[6]) {
  if (!h || !pose) return NDT_ERR_INVALID_ARGUMENT;
  h->has_reg_pose = true;
  for (int i = 0; i < 3; ++i) h->reg_t[i] = pose[i];
  return NDT_OK;
}

int ndt_hip_clear_regularization_pose(ndt_handle* h) {
  if (!h) return NDT_ERR_INVALID_ARGUMENT;
  h->has_reg_pose = false;
  return NDT_OK;
}

int ndt_hip_laplace_covariance_xy(const double hess[36], double cov[4]) {
  if (!hess || !cov) return NDT_ERR_INVALID_ARGUMENT;
  const double a = hess[0], b = hess[1], c = hess[6], d = hess[7];
  const double det = a * d - b * c;
  if (det == 0.0) {
    cov[0] = 1.0; cov[1] = 0.0; cov[2] = 0.0; cov[3] = 1.0;
    return NDT_OK;
  }
  cov[0] = -(d / det);
  cov[1] = b / det;
  cov[2] = c / det;
  cov[3] = -(a / det);
  return NDT_OK;
}

int ndt_hip_laplace_covariance_6x6(const double hess[36], double cov[36]) {
  if (!hess || !cov) return NDT_ERR_INVALID_ARGUMENT;
  // -H^{-1} by Gauss-Jordan with partial pivoting; identity if singular.
  double a[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 12; ++j) a[i][j] = j < 6 ? hess[i * 6 + j] : (j - 6 == i ? 1.0 : 0.0);
  for (int c = 0; c < 6; ++c) {
    int p = c;
    for (int r = c + 1; r < 6; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (a[p][c] == 0.0) {
      for (int i = 0; i < 36; ++i) cov[i] = (i % 7 == 0) ? 1.0 : 0.0;
      return NDT_OK;
    }
    if (p != c)
      for (int j = 0; j < 12; ++j) std::swap(a[p][j], a[c][j]);
    const double piv = a[c][c];
    for (int j = 0; j < 12; ++j) a[c][j] /= piv;
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f != 0.0)
        for (int j = 0; j < 12; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) cov[i * 6 + j] = -a[i][6 + j];
  return NDT_OK;
}

int ndt_hip_get_phase_profile(const ndt_handle* h, ndt_phase_profile* out) {
  if (!h || !out) return NDT_ERR_INVALID_ARGUMENT;
  *out = h->phase_profile;
  return NDT_OK;
}

int ndt_hip_get_timing(const ndt_handle* h, ndt_timing* out) {
  if (!h || !out) return NDT_ERR_INVALID_ARGUMENT;
  *out = h->timing;
  return NDT_OK;
}

int ndt_hip_get_tuning(const ndt_handle* h, ndt_tuning_entry* out, size_t cap, size_t* count) {
  if (!h || !count || (cap > 0 && !out)) return NDT_ERR_INVALID_ARGUMENT;
  size_t k = 0;
  for (const auto& e : h->tuning) {
    if (k < cap) {
      out[k].scan_bits = e.first.first;
      out[k].batch_bits = e.first.second;
      out[k].workgroups_per_alignment = e.second.first;
      out[k].kernel_ms = e.second.second;
    }
    ++k;
  }
  *count = k;
  return NDT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Continuous-batching server (no counterpart in the reference's API: it
// serves the multi-scan path, align_parallel_scans ndt.rs:1224-1320, as a
// stream of requests instead of one batch).

#include <atomic>
#include <deque>

struct ndt_server {
  ndt_handle* h = nullptr;
  int slots = 0, W = 0;
  size_t max_points = 0;
  ndt::ServerSlot* mb = nullptr;      // pinned host memory
  ndt::ServerControl* ctl = nullptr;  // pinned host memory
  float* host_points = nullptr;       // pinned host memory, slots x max_points x 3
  DevBuf points, ws;
  hipStream_t kstream = nullptr;
  struct Pending {
    uint64_t ticket;
    std::vector<float> xyz;
    double pose[6];
  };
  std::deque<Pending> queue;
  std::vector<uint64_t> slot_ticket;  // 0: free
  std::vector<unsigned long long> slot_seq;
  unsigned long long heartbeat = 0;
  std::chrono::steady_clock::time_point last_call{};
  double max_gap_s = 0.0;  // longest time between two host calls (diagnostics)
  uint64_t next_ticket = 1;
  int wall_clock_khz = 100000;
  bool running = false;
};

// Default idle timeout of a server launch: it ends this long after the last
// submit or poll.
constexpr double kServerIdleSeconds = 10.0;

namespace {

template <class T>
void store_release(T* p, T v) {
  __atomic_store_n(p, v, __ATOMIC_RELEASE);
}
template <class T>
T load_acquire(const T* p) {
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
}

// Tells the launch that the host is still there (resets its idle timeout).
void server_heartbeat(ndt_server* s) {
  s->heartbeat += 1;
  store_release(&s->ctl->heartbeat, s->heartbeat);
  const auto now = std::chrono::steady_clock::now();
  if (s->heartbeat > 1)
    s->max_gap_s = std::max(s->max_gap_s, std::chrono::duration<double>(now - s->last_call).count());
  s->last_call = now;
}

// The launch has ended or is ending on its idle timeout: no request posted
// from now on would be run.
bool server_exited(ndt_server* s) { return load_acquire(&s->ctl->exited) != 0u; }

int server_exited_err(ndt_server* s) {
  const ndt::ServerControl* c = s->ctl;
  const double tick_s = 1e-3 / (double)s->wall_clock_khz;
  char buf[384];
  std::snprintf(buf, sizeof(buf),
                "server kernel has exited (no submit or poll within its idle timeout): slot %llu waited %.3f s "
                "against a timeout of %.3f s at heartbeat %llu (host heartbeat %llu, longest host gap between "
                "calls %.3f s); restart the server",
                (unsigned long long)load_acquire(&c->exit_slot), (double)load_acquire(&c->exit_waited) * tick_s,
                (double)load_acquire(&c->exit_timeout) * tick_s, (unsigned long long)load_acquire(&c->exit_heartbeat),
                (unsigned long long)s->heartbeat, s->max_gap_s);
  return set_err(s->h, NDT_ERR_HIP, buf);
}

// Posts queued requests to free slots (host memory writes only: no runtime
// call that could need the device while the server occupies it).
int server_admit(ndt_server* s) {
  if (s->queue.empty()) return NDT_OK;
  if (server_exited(s)) return server_exited_err(s);
  for (int k = 0; k < s->slots && !s->queue.empty(); ++k) {
    if (s->slot_ticket[k]) continue;
    ndt_server::Pending& p = s->queue.front();
    const size_t n = p.xyz.size() / 3;
    std::memcpy(s->host_points + (size_t)k * s->max_points * 3, p.xyz.data(), n * 12);
    s->mb[k].n = (uint32_t)n;
    for (int i = 0; i < 6; ++i) s->mb[k].pose[i] = p.pose[i];
    s->slot_seq[k] += 1;
    store_release(&s->mb[k].seq, s->slot_seq[k]);
    s->slot_ticket[k] = p.ticket;
    s->queue.pop_front();
  }
  return NDT_OK;
}

}  // namespace

int ndt_hip_server_start(ndt_handle* h, int32_t slots, int32_t wg_per_slot, size_t max_points, ndt_server** out) {
  if (!h || !out || slots <= 0 || wg_per_slot <= 0 || max_points == 0 || max_points >= (1u << 31))
    return set_err(h, NDT_ERR_INVALID_ARGUMENT, "invalid server shape");
  *out = nullptr;
  if (int rc = check_no_server(h)) return rc;
  if (!h->grid.ready) return set_err(h, NDT_ERR_NO_TARGET, "No target point cloud set");
  if ((size_t)slots * wg_per_slot > (size_t)std::max(1, h->coop_capacity))
    return set_err(h, NDT_ERR_CAPACITY, "slots x workgroups exceed the co-resident workgroups");
  auto* s = new ndt_server();
  s->h = h;
  s->slots = slots;
  s->W = wg_per_slot;
  s->max_points = max_points;
  s->slot_ticket.assign(slots, 0);
  s->slot_seq.assign(slots, 0);
  auto fail = [&](hipError_t e, const char* what) {
    ndt_hip_server_stop(s);
    return hip_err(h, e, what);
  };
  hipError_t e;
  const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->mb), sizeof(ndt::ServerSlot) * slots, flags)) != hipSuccess)
    return fail(e, "server mailbox");
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->ctl), sizeof(ndt::ServerControl), flags)) != hipSuccess)
    return fail(e, "server control");
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->host_points), (size_t)slots * max_points * 12, flags)) !=
      hipSuccess)
    return fail(e, "server points");
  std::memset(s->mb, 0, sizeof(ndt::ServerSlot) * slots);
  std::memset(s->ctl, 0, sizeof(ndt::ServerControl));
  s->wall_clock_khz = h->wall_clock_khz > 0 ? h->wall_clock_khz : 100000;
  s->ctl->idle_ticks = (unsigned long long)(kServerIdleSeconds * 1e3 * (double)s->wall_clock_khz);
  if ((e = s->points.reserve((size_t)slots * max_points * 12)) != hipSuccess) return fail(e, "server points");
  if ((e = s->ws.reserve(ndt::server_workspace_bytes(slots, wg_per_slot))) != hipSuccess)
    return fail(e, "server workspace");
  if ((e = hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
  const ndt::NewtonConfig cfg = newton_config(h, wg_per_slot);
  // The map (built on the handle's stream) must be complete before the
  // launch reads it; the mailbox is host memory written above, and the
  // workspace counters are cleared on the server's own stream ahead of it.
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e, "server setup");
  if ((e = ndt::launch_server(h->grid.view, cfg, s->mb, s->ctl, slots, s->host_points,
                              reinterpret_cast<float*>(s->points.ptr), (uint32_t)max_points, s->ws.ptr, s->kstream)) !=
      hipSuccess)
    return fail(e, "server launch");
  s->running = true;
  h->server = s;
  *out = s;
  return NDT_OK;
}

int ndt_hip_server_submit(ndt_server* s, const float* xyz, size_t n, const double pose[6], uint64_t* ticket) {
  if (!s || !ticket || (!xyz && n) || !pose) return NDT_ERR_INVALID_ARGUMENT;
  ndt_handle* h = s->h;
  if (n == 0) return set_err(h, NDT_ERR_EMPTY_INPUT, "Source point cloud is empty");
  if (n > s->max_points) return set_err(h, NDT_ERR_CAPACITY, "request exceeds the server's points per slot");
  if (!s->running) return set_err(h, NDT_ERR_INVALID_ARGUMENT, "server is not running");
  server_heartbeat(s);
  if (server_exited(s)) return server_exited_err(s);
  ndt_server::Pending p;
  p.ticket = s->next_ticket++;
  p.xyz.assign(xyz, xyz + 3 * n);
  for (int i = 0; i < 6; ++i) p.pose[i] = pose[i];
  s->queue.push_back(std::move(p));
  *ticket = s->queue.back().ticket;
  return server_admit(s);
}

int ndt_hip_server_poll(ndt_server* s, ndt_result* results, uint64_t* tickets, size_t cap, size_t* count) {
  if (!s || !count || (cap && (!results || !tickets))) return NDT_ERR_INVALID_ARGUMENT;
  ndt_handle* h = s->h;
  *count = 0;
  server_heartbeat(s);
  for (int k = 0; k < s->slots && *count < cap; ++k) {
    if (!s->slot_ticket[k]) continue;
    if (load_acquire(&s->mb[k].done) != s->slot_seq[k]) continue;
    results[*count] = s->mb[k].result;
    tickets[*count] = s->slot_ticket[k];
    ++*count;
    s->slot_ticket[k] = 0;
  }
  int rc = server_admit(s);
  if (rc) return rc;
  bool busy = !s->queue.empty();
  for (int k = 0; k < s->slots; ++k) busy |= s->slot_ticket[k] != 0;
  // Liveness comes from the launch's own exited word.  A poll loop that also
  // called hipStreamQuery on the server's stream was seen to stall in it
  // until the launch ended: the stalled loop stops the heartbeat, and the
  // idle timeout then ends the launch.
  if (busy && server_exited(s)) return server_exited_err(s);
  (void)h;
  return NDT_OK;
}

int ndt_hip_server_set_idle_timeout(ndt_server* s, double seconds) {
  if (!s) return NDT_ERR_INVALID_ARGUMENT;
  if (!(seconds > 0.0) || !(seconds <= 3600.0)) return set_err(s->h, NDT_ERR_INVALID_ARGUMENT, "idle timeout");
  server_heartbeat(s);
  store_release(&s->ctl->idle_ticks, (unsigned long long)(seconds * 1e3 * (double)s->wall_clock_khz));
  return NDT_OK;
}

int ndt_hip_server_stop(ndt_server* s) {
  if (!s) return NDT_OK;
  ndt_handle* h = s->h;
  int rc = NDT_OK;
  if (s->running) {
    // Requests still on a slot finish first; queued ones are dropped.
    s->queue.clear();
    // Bounded: a slot that stops answering (the launch gone) is waited for
    // no longer than the idle timeout plus a second.
    const double wait_s = (double)load_acquire(&s->ctl->idle_ticks) * 1e-3 / (double)s->wall_clock_khz + 1.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      bool busy = false;
      for (int k = 0; k < s->slots; ++k)
        if (s->slot_ticket[k] && load_acquire(&s->mb[k].done) != s->slot_seq[k]) busy = true;
      if (!busy || server_exited(s)) break;
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > wait_s) break;
      server_heartbeat(s);
    }
    store_release(&s->ctl->stop, 1u);
    hipError_t e = hipStreamSynchronize(s->kstream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) rc = hip_err(h, e, "server kernel");
    s->running = false;
  }
  if (s->kstream) (void)hipStreamDestroy(s->kstream);
  s->points.release();
  s->ws.release();
  if (s->mb) (void)hipHostFree(s->mb);
  if (s->ctl) (void)hipHostFree(s->ctl);
  if (s->host_points) (void)hipHostFree(s->host_points);
  if (h && h->server == s) h->server = nullptr;
  delete s;
  return rc;
}
