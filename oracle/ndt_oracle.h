/*
 * CPU oracle for the NDT scan-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C++ with FP64
 * arithmetic, the reference crate's CPU path (the one selected by
 * NDT_USE_GPU=0).  It is the checker the GPU library is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product (cuda_ndt_matcher_amd) never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference repository root).
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_abi.py):
 *   oracle_gauss              closed form of GaussianParams::new
 *   oracle_point_keys         the reference's key cases and brute force at cell edges
 *   oracle_grid_build/export  numpy covariance, the <6-point drop rule, the
 *                             reference's cluster and planar cases, the f32 ratio
 *                             of the eigen threshold
 *   oracle_sym_eigen3/6       LAPACK (numpy.linalg.eigh)
 *   oracle_search             brute force over all voxels; direct-mode cell sets
 *   oracle_derivatives        score -d1 at the mean, central finite differences of
 *                             score and gradient (both metrics)
 *   oracle_newton_step        the reference's negative-definite cases, numpy solve
 *   oracle_align              the reference's alignment tests (half-cubic cloud),
 *                             best-pose rule at max iterations, line search
 *   oracle_nvtl/_tp/per_point, oracle_nvtl_cell27   brute force
 *   oracle_more_thuente       the reference's cases and the strong Wolfe conditions
 *   oracle_regularization     finite differences
 *   oracle_oscillation        the reference's cases
 *   oracle_laplace_xy         closed-form 2x2 inverse
 *   oracle_pose_to_matrix/matrix_to_pose   round trip and rotation-order cases
 *   oracle_filter_points      the reference's filter and downsampling cases
 *                             (filtering/cpu.rs tests), brute-force numpy
 *                             grouping with f64 sums in input order
 * Parity unpinned (the reference is Rust and cannot be built here, so these
 * third-party results are restated, not compared bit for bit): the rounding
 * of nalgebra 0.33's symmetric_eigen, SVD solve and Euler conversions, and
 * kiddo 4.2's radius search at the exact boundary |q - mean|^2 == r^2 (taken
 * here as excluded) and its order among equal distances.
 */
#ifndef NDT_ORACLE_H
#define NDT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_grid oracle_grid;

/* Neighbour-search modes (mirrors the product's ndt_search_mode). */
enum {
  ORACLE_SEARCH_KDTREE = 0,   /* radius = resolution, nearest first (search.rs:90-99) */
  ORACLE_SEARCH_DIRECT26 = 1, /* the 26 cells around the point's cell (pclomp getNeighborhoodAtPoint) */
  ORACLE_SEARCH_DIRECT7 = 2,  /* own cell + 6 face neighbours */
  ORACLE_SEARCH_DIRECT1 = 3,  /* own cell only */
  ORACLE_SEARCH_CELL27 = 4    /* 3x3x3 cells by floor(x*inv_res) in f64, no radius (nvtl.rs:160-190) */
};

typedef struct oracle_params {
  double resolution;      /* voxel side, also the search radius */
  double outlier_ratio;
  uint32_t max_iterations;
  double trans_epsilon;
  double step_size;
  int32_t use_line_search;
  int32_t search_mode;
  int32_t distance_metric; /* 0 = point-to-distribution, 1 = point-to-plane */
  int32_t regularization_enabled;
  double regularization_scale_factor;
  int32_t has_regularization_pose;
  double regularization_translation[3];
  int32_t fp64_voxels;     /* 1 = use the FP64 voxel statistics (no f32 rounding) */
  /* Conditioning probe (tests only): when > 0, every iteration's score is
   * scaled by 1 + noise*u and each gradient / Hessian entry moves by
   * noise*u times the largest entry of its array (u in [-1, 1) from
   * noise_seed; the Hessian stays symmetric), to measure how a run responds
   * to rounding-level differences. */
  double noise;
  uint32_t noise_seed;
} oracle_params;

typedef struct oracle_result {
  double pose[6];          /* [tx,ty,tz,roll,pitch,yaw] */
  int32_t status;          /* 0 converged, 1 max iterations, 2 no correspondences, 3 singular */
  int32_t iterations;
  double score;
  double transform_probability;
  double nvtl;
  double hessian[36];      /* row-major */
  uint32_t num_correspondences;
  uint32_t oscillation_count;
} oracle_result;

/* GaussianParams::new (src/ndt_cuda/src/derivatives/types.rs:61-75). */
void oracle_gauss(double resolution, double outlier_ratio, double* d1, double* d2);

/* VoxelCoord::from_point (voxel_grid/types.rs:206-212) for every point. */
void oracle_point_keys(const float* xyz, size_t n, float resolution, int32_t* keys);

/* build_voxel_grid_cpu + Voxel::from_statistics + regularize_covariance
 * (voxel_grid/cpu.rs:51-75, types.rs:54-190).  Voxels are exported sorted by
 * (x,y,z) coordinate; the reference's own order is HashMap iteration order. */
oracle_grid* oracle_grid_build(const float* xyz, size_t n, float resolution,
                               uint32_t min_points, float eigen_ratio);
void oracle_grid_free(oracle_grid* g);
size_t oracle_grid_size(const oracle_grid* g);
void oracle_grid_export(const oracle_grid* g, int32_t* coords, float* means, float* covs,
                        float* invs, float* axes, uint32_t* counts, double* means64,
                        double* invs64);

/* Neighbour lookup.  out_idx receives up to max_k voxel indices per query (in
 * the oracle's visiting order), out_count the full count. */
void oracle_search(const oracle_grid* g, int32_t mode, const float* q, size_t n, uint32_t* out_count,
                   int32_t* out_idx, uint32_t max_k);

/* compute_derivatives_cpu_with_metric (derivatives/cpu.rs:417-538). */
void oracle_derivatives(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                        const oracle_params* p, int32_t compute_hessian, double* score,
                        double* grad6, double* hess36, uint32_t* num_corr);

/* newton_step_negative_definite (optimization/newton.rs:160-176). Returns 1 and
 * fills delta when a step exists, 0 for the singular case. */
int32_t oracle_newton_step(const double grad6[6], const double hess36[36], double min_eigenvalue,
                           double tolerance, double delta6[6]);

/* Symmetric eigendecomposition used for voxels and the Newton step. */
void oracle_sym_eigen3(const double a[9], double evals[3], double evecs[9]);
void oracle_sym_eigen6(const double a[36], double evals[6], double evecs[36]);

/* NdtOptimizer::align (optimization/solver.rs:135-305). */
void oracle_align(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                  const oracle_params* p, oracle_result* out);

/* Scoring (scoring/nvtl.rs:211-268, scoring/metrics.rs:58-178, nvtl.rs:83-190). */
double oracle_nvtl(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                   const oracle_params* p);
double oracle_transform_probability(const oracle_grid* g, const float* xyz, size_t n,
                                    const double pose[6], const oracle_params* p);
void oracle_per_point_scores(const oracle_grid* g, const float* xyz, size_t n,
                             const double pose[6], const oracle_params* p, double* out);
double oracle_nvtl_cell27(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                          const oracle_params* p);

/* More-Thuente search (optimization/more_thuente.rs:268-420) with a callback
 * phi(alpha) -> (score, directional derivative). */
typedef void (*oracle_score_fn)(double alpha, double* score, double* deriv, void* user);
void oracle_more_thuente(oracle_score_fn fn, void* user, double initial_score,
                         double initial_derivative, double initial_step, double step_max,
                         double step_min, double mu, double nu, uint32_t max_iterations,
                         double* step_length, int32_t* converged, uint32_t* iterations,
                         double* final_value);

/* RegularizationTerm::compute_derivatives (optimization/regularization.rs:102-159). */
void oracle_regularization(const oracle_params* p, const double pose[6], uint32_t corr, double* score,
                           double grad6[6], double hess36[36]);

/* count_oscillation (optimization/oscillation.rs:50-99) over k translations. */
uint32_t oracle_oscillation(const double* translations, size_t k);

/* estimate_xy_covariance_by_laplace (cuda_ndt_matcher/src/covariance.rs:82-98). */
void oracle_laplace_xy(const double hess36[36], double cov4[4]);

/* filter_points_cpu + voxel_downsample_cpu (filtering/cpu.rs:7-100); out holds
 * 3n floats; returns the number of points written (cells in lexicographic
 * order when downsampling). */
size_t oracle_filter_points(const float* xyz, size_t n, float min_distance, float max_distance, float min_z,
                            float max_z, int32_t downsample, float resolution, float* out, uint64_t counts[4]);

/* Isometry helpers restated from nalgebra 0.33 (used by optimization/types.rs:156-174). */
void oracle_pose_to_matrix(const double pose[6], double m12[12]);
void oracle_matrix_to_pose(const double m12[12], double pose[6]);

#ifdef __cplusplus
}
#endif
#endif
