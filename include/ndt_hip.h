/*
 * ndt_hip.h - C ABI of the MI355X NDT scan-matching library (libndt_hip.so).
 *
 * The library replaces the scan-matching hot path of the reference crate
 * `ndt_cuda` (jerry73204/cuda_ndt_matcher, src/ndt_cuda).  Each entry point
 * below names the reference interface it stands in for (file:line, relative
 * to the reference repository root).  Semantics follow the reference's CPU
 * path (NDT_USE_GPU=0), which is the one the node's results are defined by.
 *
 * Conventions
 *  - Every function returns an int: 0 on success, a negative ndt_status
 *    otherwise; ndt_hip_last_error(h) gives a message for the last failure.
 *  - Points are AoS float triples [[x,y,z],...] as in the reference's
 *    `&[[f32; 3]]`.  Host arrays are borrowed for the duration of the call.
 *    `*_device` variants take device pointers (HBM-resident inputs); the
 *    library works on its own non-blocking stream, so such inputs must be
 *    completely written (their producer's stream synchronised) before the
 *    call.
 *  - A pose is the internal 6-vector [tx,ty,tz,roll,pitch,yaw] (f64) that the
 *    reference produces with isometry_to_pose_vector
 *    (src/ndt_cuda/src/optimization/types.rs:164-174, nalgebra euler_angles)
 *    and turns back into an Isometry3 with pose_vector_to_isometry
 *    (types.rs:156-161).  Callers holding an Isometry3 apply those two
 *    conversions on their side of the boundary.
 *  - A handle owns one HIP stream and all of its device memory.  A handle is
 *    not thread-safe; different handles may be used concurrently.
 */
#ifndef NDT_HIP_H
#define NDT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes.  Mirrors the reference's error classes: anyhow bails in
 * ndt.rs (empty cloud, no target, zero voxels) and CudaError from cuda_ffi
 * (src/cuda_ffi/src/radix_sort.rs:59-94). */
typedef enum ndt_status {
  NDT_OK = 0,
  NDT_ERR_INVALID_ARGUMENT = -1, /* null pointer, bad enum, bad size */
  NDT_ERR_EMPTY_INPUT = -2,      /* "Target point cloud is empty" / "Source point cloud is empty" */
  NDT_ERR_NO_TARGET = -3,        /* "No target set. Call set_target() first." */
  NDT_ERR_NO_VOXELS = -4,        /* "No voxels created from target points (too sparse?)" */
  NDT_ERR_HIP = -5,              /* a HIP runtime call failed */
  NDT_ERR_UNSUPPORTED = -6,      /* a configuration this build does not implement */
  NDT_ERR_CAPACITY = -7          /* input exceeds a documented limit */
} ndt_status;

/* Neighbour search for the derivative, NVTL and TP passes. */
typedef enum ndt_search_mode {
  NDT_SEARCH_KDTREE = 0,   /* every voxel whose mean is within `resolution` (the reference's
                              KD-tree radius search, voxel_grid/search.rs:90-99) */
  NDT_SEARCH_DIRECT26 = 1, /* the 26 cells around the point's cell (Autoware pclomp DIRECT26) */
  NDT_SEARCH_DIRECT7 = 2,  /* the point's cell and its 6 face neighbours */
  NDT_SEARCH_DIRECT1 = 3   /* the point's cell */
} ndt_search_mode;

typedef enum ndt_distance_metric {
  NDT_METRIC_POINT_TO_DISTRIBUTION = 0, /* derivatives/types.rs:10-25 */
  NDT_METRIC_POINT_TO_PLANE = 1
} ndt_distance_metric;

typedef enum ndt_precision {
  NDT_PRECISION_FP64 = 0,    /* per-point terms and accumulation in f64 */
  NDT_PRECISION_FP32 = 1     /* per-point terms in f32 with compensated (two-sum) f64-free
                                accumulation; f64 for the 6x6 solve and the pose */
} ndt_precision;

/* NdtScanMatcherConfig (src/ndt_cuda/src/ndt.rs:41-107) plus the build's own
 * knobs.  ndt_hip_config_default() fills the reference defaults. */
typedef struct ndt_config {
  float resolution;                   /* 2.0 */
  uint32_t max_iterations;            /* 30 */
  double trans_epsilon;               /* 0.01 */
  double step_size;                   /* 0.1 */
  double outlier_ratio;               /* 0.55 */
  double regularization;              /* 0.001 (kept for API parity; the CPU path ignores it) */
  int32_t use_line_search;            /* 0 */
  uint32_t min_points_per_voxel;      /* 6 */
  int32_t regularization_enabled;     /* 0: GNSS regularisation term (optimization/regularization.rs) */
  double regularization_scale_factor; /* 0.01 */
  int32_t distance_metric;            /* ndt_distance_metric */
  int32_t search_mode;                /* ndt_search_mode, default NDT_SEARCH_KDTREE */
  int32_t precision;                  /* ndt_precision, default NDT_PRECISION_FP64 */
  int32_t workgroups_per_alignment;   /* 0 = automatic; otherwise workgroups cooperating on one alignment */
  int32_t profile_phases;             /* 1 = time the Newton kernel's phases (see ndt_hip_get_phase_profile) */
  int32_t cell_tables;                /* 1 = per-cell neighbour tables when the map's padded voxel box has
                                         at most 2^26 cells (20 bytes per cell); 0 = hash + bitmap lookups only.
                                         Both give the same neighbours in the same order. */
  int32_t use_graph;                  /* 1 = align* calls replay a HIP graph per launch shape (jobs copy, counter
                                         fill, persistent Newton kernel, results copy), rebuilt only when the
                                         shape or a buffer changes; 0 = direct launches.  Same results. */
  int32_t autotune;                   /* 1 = with workgroups_per_alignment 0, the first align* call of each
                                         (scan size, batch size) bucket times candidate workgroup counts and
                                         keeps the fastest for that bucket (see ndt_hip_get_tuning). */
} ndt_config;

/* Convergence status (optimization/types.rs:66-82). */
typedef enum ndt_convergence {
  NDT_CONVERGED = 0,
  NDT_MAX_ITERATIONS = 1,
  NDT_NO_CORRESPONDENCES = 2,
  NDT_SINGULAR_HESSIAN = 3
} ndt_convergence;

/* AlignResult (ndt.rs:220-248) with NdtResult's status (optimization/types.rs:100-131). */
typedef struct ndt_result {
  double pose[6];
  int32_t status;                /* ndt_convergence */
  int32_t converged;             /* status == NDT_CONVERGED */
  double score;
  double transform_probability;  /* score / number of source points (solver.rs:707-713) */
  double nvtl;
  uint32_t iterations;
  uint32_t num_correspondences;
  uint32_t oscillation_count;
  uint32_t pair_evaluations;     /* point-voxel pairs summed over the derivative passes with the Hessian
                                    (work done; not part of the reference's result) */
  double hessian[36];            /* row-major */
} ndt_result;

/* Device timing of the last call on a handle, from HIP events recorded on the
 * handle's stream around each kernel group. */
typedef struct ndt_timing {
  float build_ms;     /* last set_target: whole voxel-grid build */
  float sort_ms;      /* part of build_ms spent in the radix sort passes */
  float align_ms;     /* last align*: the persistent Newton kernel */
  float score_ms;     /* last evaluate_*: the scoring kernel */
  uint32_t align_workgroups;
  uint32_t align_slots;
  uint32_t align_graph;  /* 1 when the last align* replayed a graph */
  uint32_t align_tuned;  /* 1 when its workgroup count came from the tuning table */
} ndt_timing;

/* Per-phase device time of the last align* call, from wall-clock stamps taken
 * by workgroup 0 of slot 0 inside the Newton kernel (profile_phases = 1).
 * Times are totals over the launch in microseconds. */
typedef struct ndt_phase_profile {
  float derivative_us;  /* per-point derivative passes (this workgroup's share) + block reduction */
  float exchange_us;    /* partial exchange and slot barrier (0 with one workgroup per slot) */
  float solve_us;       /* 6x6 Newton step (waves 0 and 1), convergence test and pose update */
  float line_search_us; /* More-Thuente evaluations (derivative passes without Hessian) */
  float nvtl_us;        /* final NVTL pass */
  float total_us;
  uint32_t iterations;
  uint32_t line_search_evaluations;
  float slowest_derivative_us; /* derivative passes of the slowest workgroup of the launch */
  float solve_setup_us;        /* part of solve_us before the 6x6 step (sums, best pose, regularisation) */
  float solve_step_us;         /* part of solve_us in the 6x6 step itself */
  uint32_t eigen_solves;       /* steps that needed the eigendecomposition (Hessian not negative definite) */
  uint32_t jacobi_sweeps;      /* Jacobi sweeps over those steps */
} ndt_phase_profile;

typedef struct ndt_handle ndt_handle;

void ndt_hip_config_default(ndt_config* cfg);

/* NdtScanMatcher::with_config / builder().build() (ndt.rs:287-330, 200-210). */
int ndt_hip_create(const ndt_config* cfg, ndt_handle** out);
void ndt_hip_destroy(ndt_handle* h);
const char* ndt_hip_last_error(const ndt_handle* h);

/* NdtScanMatcher::set_target (ndt.rs:361-415).  Builds the voxel grid on the
 * GPU (Morton keys, LDS-staged radix sort, segmented moments, eigen
 * regularisation, cell tables or voxel hash).  Replaces the CUB bindings of cuda_ffi
 * (radix_sort.cu:24,62; segment_detect.cu; voxel_hash.cu:218-386). */
int ndt_hip_set_target(ndt_handle* h, const float* xyz, size_t n);
int ndt_hip_set_target_device(ndt_handle* h, const float* d_xyz, size_t n);
/* has_target / target_grid().len() (ndt.rs:343-355). */
int ndt_hip_has_target(const ndt_handle* h);
int ndt_hip_num_voxels(const ndt_handle* h, size_t* out);

/* Copies the voxel grid to host arrays (any pointer may be NULL).  Voxels are
 * ordered by Morton key.  Sizes: coords 3*V, means 3*V, covs 9*V, invs 9*V,
 * axes 3*V, counts V. */
int ndt_hip_get_voxels(ndt_handle* h, int32_t* coords, float* means, float* covs, float* invs, float* axes,
                       uint32_t* counts);

/* Build internals for parity tests: for every map point its voxel key (3*N
 * int32), the stable sort permutation (N uint32: original index at each sorted
 * position) and the packed Morton key at each sorted position (N uint64). */
int ndt_hip_get_build_debug(ndt_handle* h, int32_t* point_keys, uint32_t* sorted_index, uint64_t* sorted_keys);

/* Neighbour lookup for parity tests: for each query point (f32 xyz), up to
 * max_k voxel indices (into the ndt_hip_get_voxels order) and the full count. */
int ndt_hip_lookup(ndt_handle* h, int32_t mode, const float* q, size_t n, uint32_t max_k, uint32_t* counts,
                   int32_t* indices);

/* compute_derivatives_cpu_with_metric (derivatives/cpu.rs:417-538): score,
 * gradient, Hessian and correspondence count at one pose. */
int ndt_hip_derivatives(ndt_handle* h, const float* xyz, size_t n, const double pose[6], int32_t compute_hessian,
                        double* score, double grad[6], double hess[36], uint32_t* num_correspondences);

/* NdtScanMatcher::align (ndt.rs:421-476) -> NdtOptimizer::align (solver.rs:135-305). */
int ndt_hip_align(ndt_handle* h, const float* xyz, size_t n, const double pose[6], ndt_result* out);
int ndt_hip_align_device(ndt_handle* h, const float* d_xyz, size_t n, const double pose[6], ndt_result* out);

/* Allocation-free single alignment.  No counterpart in the reference: it is
 * for callers that cannot allocate device or pinned memory during a call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same computation and result as ndt_hip_align_device with the
 * same workgroup count (the autotuner and phase stamps are not used), but
 * the result is written to d_out in device memory, and the job and exchange
 * state live in a fixed device array of about 131 KiB per GPU, allocated
 * with the library and held until the process ends: at most 256 workgroups
 * per alignment (NDT_ERR_CAPACITY beyond).  The call allocates nothing and
 * waits for the alignment on the handle's stream.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two threads, or two
 * handles on different streams) race and must be serialised by the caller. */
int ndt_hip_align_device_noalloc(ndt_handle* h, const float* d_xyz, size_t n, const double pose[6],
                                 ndt_result* d_out);

/* align_batch / align_batch_gpu (ndt.rs:903-1001): one scan, m initial poses,
 * all in one persistent launch. */
int ndt_hip_align_batch(ndt_handle* h, const float* xyz, size_t n, const double* poses, size_t m, ndt_result* out);
int ndt_hip_align_batch_device(ndt_handle* h, const float* d_xyz, size_t n, const double* poses, size_t m,
                               ndt_result* out);

/* align_parallel_scans (ndt.rs:1224-1320): m scans of uneven size with their
 * own poses, in one launch.  Unlike the reference's GPU batch (which reports
 * NVTL 0 and TP = score/#corr, ndt.rs:1305-1310) every result carries the
 * CPU-path NVTL and TP. */
int ndt_hip_align_scans(ndt_handle* h, const float* const* xyz, const size_t* n, const double* poses, size_t m,
                        ndt_result* out);

/* evaluate_nvtl / evaluate_transform_probability (ndt.rs:740-797). */
int ndt_hip_evaluate_nvtl(ndt_handle* h, const float* xyz, size_t n, const double pose[6], double* out);
int ndt_hip_evaluate_transform_probability(ndt_handle* h, const float* xyz, size_t n, const double pose[6],
                                           double* out);
/* compute_per_point_scores (ndt.rs:800-817, scoring/metrics.rs:137-178). */
int ndt_hip_per_point_scores(ndt_handle* h, const float* xyz, size_t n, const double pose[6], double* out);
/* evaluate_nvtl_batch (ndt.rs:833-890): m poses, 3x3x3-cell NVTL (scoring/nvtl.rs:83-190). */
int ndt_hip_evaluate_nvtl_batch(ndt_handle* h, const float* xyz, size_t n, const double* poses, size_t m,
                                double* out);

/* set_/clear_regularization_pose (ndt.rs:1151-1160); translation only is used. */
int ndt_hip_set_regularization_pose(ndt_handle* h, const double pose[6]);
int ndt_hip_clear_regularization_pose(ndt_handle* h);

/* estimate_xy_covariance_by_laplace (cuda_ndt_matcher/src/covariance.rs:82-98)
 * and its 6x6 extension -H^-1 (identity when H is singular). */
int ndt_hip_laplace_covariance_xy(const double hess[36], double cov[4]);
int ndt_hip_laplace_covariance_6x6(const double hess[36], double cov[36]);

/* Sensor point filter: FilterParams (src/ndt_cuda/src/filtering/mod.rs:13-31,
 * defaults :22-31).  min/max distance from the sensor origin, min/max z, and
 * an optional voxel-grid downsampling (downsample = 0 is the reference's
 * None).  Replaces GpuPointFilter::filter / filter_points_cpu
 * (filtering/mod.rs:61-230, cpu.rs:7-58) as called by
 * cuda_ndt_matcher/src/pointcloud.rs:59-95 (filter_sensor_points). */
typedef struct ndt_filter_params {
  float min_distance;
  float max_distance;
  float min_z;
  float max_z;
  float downsample_resolution;
  int32_t downsample;
} ndt_filter_params;

/* FilterResult's counts (filtering/mod.rs:34-41) plus the points kept by the
 * distance and z tests (before downsampling). */
typedef struct ndt_filter_counts {
  uint64_t kept;
  uint64_t removed_by_distance;
  uint64_t removed_by_z;
  uint64_t removed_by_downsampling;
} ndt_filter_counts;

void ndt_hip_filter_params_default(ndt_filter_params* p);

/* filter_points_cpu (filtering/cpu.rs:7-58) on the GPU.  Each removed point is
 * counted under the first test it fails (distance, then z; a NaN coordinate
 * passes both, as in the reference's CPU path).  Without downsampling the kept
 * points come out in input order; with it, one centroid per occupied cell of
 * floor(p * (1/res)) (f64 sums in input order, rounded to f32), in the cells'
 * key order (the reference returns HashMap order).  out (3*out_cap floats)
 * must hold the result: a smaller out_cap returns NDT_ERR_INVALID_ARGUMENT with
 * the needed count in *out_n; n >= 2^30, or cells beyond 2^20 from the origin
 * on some axis, return NDT_ERR_CAPACITY.  Needs no target; the handle's map is
 * untouched.
 * The _device form reads d_xyz and writes d_out in device memory (a filtered
 * scan can go straight to ndt_hip_align_device). */
int ndt_hip_filter_points(ndt_handle* h, const float* xyz, size_t n, const ndt_filter_params* p, float* out,
                          size_t out_cap, size_t* out_n, ndt_filter_counts* counts);
int ndt_hip_filter_points_device(ndt_handle* h, const float* d_xyz, size_t n, const ndt_filter_params* p,
                                 float* d_out, size_t out_cap, size_t* out_n, ndt_filter_counts* counts);
/* voxel_downsample_cpu (filtering/cpu.rs:67-100): resolution <= 0 returns the
 * points unchanged. */
int ndt_hip_voxel_downsample(ndt_handle* h, const float* xyz, size_t n, float resolution, float* out, size_t out_cap,
                             size_t* out_n);

/* Continuous-batching server.  No counterpart in the reference: it serves
 * the multi-scan path (align_parallel_scans, ndt.rs:1224-1320) as a stream of
 * requests.  One persistent launch holds `slots` slots of `wg_per_slot`
 * workgroups; a request is posted to a free slot as soon as one frees up, so
 * requests that arrive while others run are admitted without waiting for a
 * batch to end.  submit copies the points (host) and queues the request;
 * poll collects finished results (up to cap, with their tickets) and posts
 * queued requests to free slots; stop lets running requests finish, drops
 * queued ones and ends the launch.  Results equal an align_scans /
 * ndt_hip_align of the same scan with workgroups_per_alignment = wg_per_slot.
 * While a server runs it holds the device's co-resident workgroups, so the
 * handle's other calls (and a second start) return NDT_ERR_INVALID_ARGUMENT
 * until stop; destroying the handle stops it.  The whole launch ends when no
 * submit or poll has been made for the idle timeout (10 s by default): every
 * slot leaves, submit and poll then return NDT_ERR_HIP, and the server must
 * be stopped and started again.  A caller that waits for requests keeps the
 * launch alive by polling, and a slot that finishes a job keeps the idle
 * slots alive.  hipFree waits for every launch on the device to end: while
 * a server runs, freeing device memory anywhere in the process (destroying
 * another handle, a call that grows a handle's buffers) blocks until the
 * server ends, so such calls belong before start or after stop. */
typedef struct ndt_server ndt_server;
int ndt_hip_server_start(ndt_handle* h, int32_t slots, int32_t wg_per_slot, size_t max_points, ndt_server** out);
int ndt_hip_server_submit(ndt_server* s, const float* xyz, size_t n, const double pose[6], uint64_t* ticket);
int ndt_hip_server_poll(ndt_server* s, ndt_result* results, uint64_t* tickets, size_t cap, size_t* count);
/* Sets the idle timeout (0 < seconds <= 3600) of a running server; counts
 * from this call. */
int ndt_hip_server_set_idle_timeout(ndt_server* s, double seconds);
int ndt_hip_server_stop(ndt_server* s);

/* Learned initial-pose proposal.  No counterpart in the reference: it seeds
 * the particle search of estimate_initial_pose (cuda_ndt_matcher/src/
 * initial_pose.rs:115-400) with hypotheses from a multilayer perceptron over
 * a bird's-eye-view descriptor of the scan.  Training and inference run on
 * bf16 MFMA kernels with FP32 accumulation and FP32 master weights (Adam).
 *
 * sizes[0..n_sizes): layer widths, each a multiple of 32; hidden layers use
 * GELU, the last is linear.  Parameters (get/set/gradients) are laid out per
 * layer as W^T (out x in, row-major, f32) followed by the bias (out).  Batch
 * sizes are multiples of 32.  train_step: one Adam step on the
 * winner-takes-all proposal loss; the output holds H = out / 4 hypotheses
 * (x, y, yaw, logit) and labels are (x, y, yaw) per sample. */
typedef struct ndt_mlp ndt_mlp;
int ndt_hip_mlp_create(const int32_t* sizes, int32_t n_sizes, uint64_t seed, ndt_mlp** out);
void ndt_hip_mlp_destroy(ndt_mlp* m);
const char* ndt_hip_mlp_last_error(const ndt_mlp* m);
size_t ndt_hip_mlp_num_params(const ndt_mlp* m);
int ndt_hip_mlp_get_params(ndt_mlp* m, float* params);
int ndt_hip_mlp_set_params(ndt_mlp* m, const float* params);
/* Forward of bf16 inputs in device memory (batch x sizes[0]); out: host, batch x sizes[last].
 * The network runs on its own stream, which is not ordered after other
 * streams: d_x must be completely written before the call (as for the
 * training step's inputs). */
int ndt_hip_mlp_forward_device(ndt_mlp* m, const uint16_t* d_x, size_t batch, float* out);
/* Forward and backward of host inputs (rounded to bf16) with a given output
 * gradient dout; returns the outputs and the parameter gradients. */
int ndt_hip_mlp_forward_backward(ndt_mlp* m, const float* x, size_t batch, const float* dout, float* out,
                                 float* grads);
int ndt_hip_mlp_train_step_device(ndt_mlp* m, const uint16_t* d_x, const float* labels, size_t batch,
                                  float learning_rate, float* mean_loss);
/* Bird's-eye-view descriptors, one per output row j.  Row j describes scan
 * q = d_index ? d_index[j] : j, whose points are d_xyz[offsets[q] ..
 * offsets[q+1]) in the sensor frame.  When d_rot is not null the points are
 * first rotated about z by (c, s) = (d_rot[2j], d_rot[2j+1]) (8-byte
 * aligned): x' = c x - s y, y' = s x + c y, each product and difference
 * rounded to f32 on its own.  A scan at map yaw t rotated by d reads as the
 * same place at yaw t - d, which is how training re-labels a scan at many
 * headings.  Output: log(1 + count) / 4 per cell of an nx x ny x nz grid
 * (x/y cells of `cell` metres centred on the sensor, heights from zmin in
 * steps of zcell), bf16, cell index (iy * nx + ix) * nz + iz.
 * nx * ny * nz <= 32768. */
int ndt_hip_bev_device(const float* d_xyz, const uint32_t* d_offsets, const uint32_t* d_index, size_t rows,
                       const float* d_rot, int32_t nx, int32_t ny, int32_t nz, float cell, float zmin, float zcell,
                       uint16_t* d_bev);

int ndt_hip_get_timing(const ndt_handle* h, ndt_timing* out);

/* Autotuner table (config autotune = 1; no counterpart in the reference):
 * one entry per (scan size, batch size) bucket seen so far.  scan_bits is
 * the bit length of the largest scan's point count, batch_bits that of the
 * number of alignments in the call.  *count receives the number of entries;
 * at most cap are written. */
typedef struct ndt_tuning_entry {
  int32_t scan_bits;
  int32_t batch_bits;
  int32_t workgroups_per_alignment;
  float kernel_ms; /* best of two timed launches at that count */
} ndt_tuning_entry;
int ndt_hip_get_tuning(const ndt_handle* h, ndt_tuning_entry* out, size_t cap, size_t* count);
int ndt_hip_get_phase_profile(const ndt_handle* h, ndt_phase_profile* out);

#ifdef __cplusplus
}
#endif
#endif
