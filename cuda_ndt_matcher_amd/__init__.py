"""MI355X-native NDT scan matching (the ndt_cuda hot path on gfx950).

The compute lives in libndt_hip.so (hand-written HIP kernels, C ABI in
include/ndt_hip.h); this package is the host-side mirror of the reference
crate's NdtScanMatcher API plus seeded synthetic workloads.
"""
from .dual import DualMatcher, UpdateStatus
from .filtering import FilterParams, FilterResult, GpuPointFilter, PointFilterParams, filter_sensor_points
from .initial_pose import InitialPoseResult, Particle, TreeStructuredParzenEstimator, estimate_initial_pose
from .isometry import Isometry3
from .matcher import (METRIC_POINT_TO_DISTRIBUTION, METRIC_POINT_TO_PLANE, PRECISION_FP32, PRECISION_FP64,
                      SEARCH_DIRECT1, SEARCH_DIRECT7, SEARCH_DIRECT26, SEARCH_KDTREE, AlignResult, NdtError,
                      AlignServer, NdtScanMatcher, NdtScanMatcherBuilder, NdtScanMatcherConfig, laplace_covariance_6x6,
                      laplace_covariance_xy)

__all__ = [
    "Isometry3", "AlignResult", "NdtError", "NdtScanMatcher", "NdtScanMatcherBuilder", "NdtScanMatcherConfig",
    "laplace_covariance_xy", "laplace_covariance_6x6", "SEARCH_KDTREE", "SEARCH_DIRECT26", "SEARCH_DIRECT7",
    "SEARCH_DIRECT1", "METRIC_POINT_TO_DISTRIBUTION", "METRIC_POINT_TO_PLANE", "PRECISION_FP64", "PRECISION_FP32",
    "estimate_initial_pose", "InitialPoseResult", "Particle", "TreeStructuredParzenEstimator",
    "DualMatcher", "UpdateStatus", "FilterParams", "PointFilterParams", "FilterResult", "GpuPointFilter",
    "filter_sensor_points",
]
