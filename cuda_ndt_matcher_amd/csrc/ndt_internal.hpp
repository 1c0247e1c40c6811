// Host-side state shared by the build, the Newton/scoring launchers and the
// C ABI.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/ndt_hip.h"
#include "ndt_device.hpp"

namespace ndt {

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t want) {
    if (want <= bytes && ptr) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    size_t cap = want < 256 ? 256 : want;
    hipError_t e = hipMalloc(&ptr, cap);
    if (e == hipSuccess) bytes = cap;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

struct DeviceGrid {
  // inputs
  const float* points = nullptr;  // device, 3*n_points
  DevBuf owned_points;
  size_t n_points = 0;
  // sort scratch
  DevBuf sort_keys_a, sort_keys_b, sort_vals_a, sort_vals_b, hist, tile_hist, tile_heads, seg_start, scalars;
  void* sorted_keys = nullptr;
  uint32_t* sorted_vals = nullptr;
  int key_bits = 0;
  bool key64 = false;
  int32_t key_min[3] = {0, 0, 0};
  int key_bits_axis[3] = {0, 0, 0};
  // voxel slots (one per segment, in Morton order)
  uint32_t segments = 0;
  uint32_t seg_cap = 0;  // capacity of vox, cov, axis, count, coord and valid, in voxels (grow-only)
  uint32_t num_valid = 0;
  DevBuf vox, cov, axis, count, coord, valid, table, bitmap, occ, cell_nbr, cell_ent;
  bool occ_zero = false;  // occ holds zeros (bitmap_pack clears what occ_set set)
  GridView view{};
  // host mirrors
  uint32_t* host_scalars = nullptr;  // pinned, 64 words
  int32_t* host_part = nullptr;      // per-workgroup bounds partials (8 words each), written by the
                                     // kernels into pinned, coherent, mapped host memory
  uint32_t* host_valid = nullptr;    // per-workgroup valid-voxel counts (occ_set), pinned and mapped
  uint32_t host_valid_cap = 0;       // entries of host_valid
  hipEvent_t ev_build0{}, ev_build1{}, ev_sort0{}, ev_sort1{};
  float build_ms = 0.f, sort_ms = 0.f;
  bool ready = false;
};

int build_grid(DeviceGrid& G, size_t n, float res, uint32_t min_points, bool cell_tables, hipStream_t st,
               std::string* err);

int filter_points(DeviceGrid& G, size_t n, float min_distance, float max_distance, float min_z, float max_z,
                  int downsample, float resolution, float* d_out, size_t out_cap, size_t* out_n, uint64_t counts[4],
                  hipStream_t st, std::string* err);

// One alignment request inside a persistent launch.
struct SlotJob {
  const float* pts;  // device AoS
  uint32_t n;
  uint32_t pad;
  double pose[6];
};

struct NewtonConfig {
  double d1, d2;
  double trans_eps;
  double step_size;
  double reg_scale;
  double reg_t[3];
  uint32_t max_iter;
  int32_t search_mode;
  int32_t metric;
  int32_t use_ls;
  int32_t reg_active;  // regularisation enabled and a reference pose set
  int32_t precision;
  int32_t wg_per_slot;
  // Slot exchange buffers written once per launch (0 = two alternating
  // buffers with arrival counters; see slot_reduce in newton.hip).
  int32_t exchanges;
  int32_t pad0;
  // Phase stamps (null = off): workgroup 0 accumulates wall-clock ticks per
  // phase into phase[0..5] and counts into phase[6..7].
  unsigned long long* phase;
};

static_assert(sizeof(ndt_result) == 6 * 8 + 8 + 3 * 8 + 16 + 36 * 8, "ndt_result layout");

// Launches the persistent Newton kernel over `slots` jobs (device array) and
// writes ndt_result per slot (device array).  ws must hold
// newton_workspace_bytes(slots, wg_per_slot).  ev_start (optional) is
// recorded after the counter reset, just before the Newton kernel.
size_t newton_workspace_bytes(int slots, int wg_per_slot, int exchanges = 0);
// Exchanges one alignment of at most max_iter iterations can make: one per
// derivative pass with the Hessian, up to 11 per iteration for the
// More-Thuente trials, one for the final NVTL.
inline int newton_exchanges(uint32_t max_iter, bool line_search) {
  return (int)max_iter * (line_search ? 12 : 1) + 1;
}
hipError_t launch_newton(const GridView& g, const NewtonConfig& cfg, const SlotJob* d_jobs, int slots,
                         ndt_result* d_out, void* ws, hipStream_t st, bool cooperative, hipEvent_t ev_start);
// The launch_newton sequence as an instantiated graph: jobs copied from
// pinned host h_jobs, counters filled by a kernel, ev0 recorded, the Newton
// kernel (cooperative when wg_per_slot > 1), ev1 recorded, results copied to
// pinned host h_out.  Every argument is fixed in the graph; replays read the
// current contents of h_jobs.
hipError_t build_newton_graph(const GridView& g, const NewtonConfig& cfg, SlotJob* d_jobs, const SlotJob* h_jobs,
                              int slots, ndt_result* d_out, ndt_result* h_out, void* ws, hipEvent_t ev0,
                              hipEvent_t ev1, hipGraph_t* graph, hipGraphExec_t* exec);

// Continuous-batching server.  The mailbox and the request points live in
// pinned host memory: the host posts a job to a free slot by writing the
// points, n and pose and then seq = previous + 1 with release semantics; the
// slot's workgroups copy the points to device memory, run the alignment, and
// workgroup 0 writes result and then done = seq.  The host sets stop only
// when every slot is idle.
struct alignas(128) ServerSlot {
  unsigned long long seq;   // host: number of jobs posted to this slot
  unsigned long long done;  // device: number of jobs finished
  uint32_t n;
  uint32_t pad;
  double pose[6];
  ndt_result result;
};
// Launch-wide control words, in pinned host memory.  The whole launch ends
// together: on stop, or when for idle_ticks of the device wall clock the host
// has not touched the server (no submit or poll, which bump heartbeat) and no
// slot has finished a job.
// The first slot that sees the idle timeout sets exited; every other slot
// leaves when it next looks for a job, and the host refuses new requests
// once it reads exited.
struct alignas(128) ServerControl {
  uint32_t stop;                 // host -> device
  uint32_t exited;               // device -> host
  unsigned long long heartbeat;  // host -> device: bumped by every submit and poll
  unsigned long long idle_ticks; // host -> device: idle timeout in wall-clock ticks
  // device -> host, written before exited by the slot that timed out: its
  // index, the wall-clock ticks it waited, the timeout it applied and the
  // heartbeat it last saw.
  unsigned long long exit_slot, exit_waited, exit_timeout, exit_heartbeat;
};
size_t server_workspace_bytes(int slots, int wg_per_slot);
// Launches the persistent server kernel (cooperative): slots x wg_per_slot
// workgroups.  host_points (pinned) and slot_points (device) each hold
// slots x max_points x 3 floats.  A slot's workgroup 0 decides for its slot
// (run the next job or leave); the slot's other workgroups follow it.
hipError_t launch_server(const GridView& g, const NewtonConfig& cfg, ServerSlot* mailbox, ServerControl* ctl,
                         int slots, const float* host_points, float* slot_points, uint32_t max_points, void* ws,
                         hipStream_t st);
int newton_max_coresident_workgroups();

// The allocation-free single alignment (ndt_hip_align_device_noalloc) keeps
// its job and exchange state in a fixed device array: one SlotJob at offset 0
// and, from offset kNoAllocWorkspaceOffset, the workspace of one slot of at
// most kNoAllocMaxWorkgroups workgroups (partials, then arrival counters).
constexpr int kNoAllocMaxWorkgroups = 256;
constexpr size_t kNoAllocWorkspaceOffset = 256;
constexpr size_t kNoAllocWorkspaceBytes =
    2ull * kNoAllocMaxWorkgroups * 32 * sizeof(double) + (1 + kNoAllocMaxWorkgroups / 16) * 32 * sizeof(uint32_t) + 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocWorkspaceOffset + kNoAllocWorkspaceBytes;
// Writes job into d_job (device) with a one-thread kernel on st, so the job
// needs no host staging buffer or runtime copy.
hipError_t store_job(const SlotJob& job, SlotJob* d_job, hipStream_t st);

// One derivative pass (score, gradient, Hessian, count) at a pose.
hipError_t launch_derivatives(const GridView& g, const NewtonConfig& cfg, const float* d_pts, uint32_t n,
                              const double pose[6], int hess, double* d_partials, int blocks, double* d_out,
                              hipStream_t st);

enum ScoreKind { kScoreNvtl = 0, kScoreTp = 1, kScorePerPoint = 2, kScoreNvtlCell27 = 3 };
// Scores `m` poses over n points.  For NVTL/TP writes out[m]; per-point writes out[n].
hipError_t launch_score(const GridView& g, double d1, double d2, int kind, const float* d_pts, uint32_t n,
                        const double* d_poses, int m, double* d_partials, int blocks_per_pose, double* d_out,
                        hipStream_t st);

hipError_t launch_lookup(const GridView& g, int mode, const float* d_q, uint32_t n, uint32_t max_k,
                         uint32_t* d_counts, int32_t* d_idx, hipStream_t st);

}  // namespace ndt
