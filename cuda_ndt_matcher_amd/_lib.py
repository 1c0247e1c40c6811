"""ctypes binding of libndt_hip.so (the C ABI in include/ndt_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libndt_hip.so")

NDT_OK = 0
ERROR_NAMES = {
    -1: "INVALID_ARGUMENT",
    -2: "EMPTY_INPUT",
    -3: "NO_TARGET",
    -4: "NO_VOXELS",
    -5: "HIP",
    -6: "UNSUPPORTED",
    -7: "CAPACITY",
}


class NdtConfig(C.Structure):
    _fields_ = [
        ("resolution", C.c_float),
        ("max_iterations", C.c_uint32),
        ("trans_epsilon", C.c_double),
        ("step_size", C.c_double),
        ("outlier_ratio", C.c_double),
        ("regularization", C.c_double),
        ("use_line_search", C.c_int32),
        ("min_points_per_voxel", C.c_uint32),
        ("regularization_enabled", C.c_int32),
        ("regularization_scale_factor", C.c_double),
        ("distance_metric", C.c_int32),
        ("search_mode", C.c_int32),
        ("precision", C.c_int32),
        ("workgroups_per_alignment", C.c_int32),
        ("profile_phases", C.c_int32),
        ("cell_tables", C.c_int32),
        ("use_graph", C.c_int32),
        ("autotune", C.c_int32),
    ]


class NdtResult(C.Structure):
    _fields_ = [
        ("pose", C.c_double * 6),
        ("status", C.c_int32),
        ("converged", C.c_int32),
        ("score", C.c_double),
        ("transform_probability", C.c_double),
        ("nvtl", C.c_double),
        ("iterations", C.c_uint32),
        ("num_correspondences", C.c_uint32),
        ("oscillation_count", C.c_uint32),
        ("pair_evaluations", C.c_uint32),
        ("hessian", C.c_double * 36),
    ]


class NdtTiming(C.Structure):
    _fields_ = [
        ("build_ms", C.c_float),
        ("sort_ms", C.c_float),
        ("align_ms", C.c_float),
        ("score_ms", C.c_float),
        ("align_workgroups", C.c_uint32),
        ("align_slots", C.c_uint32),
        ("align_graph", C.c_uint32),
        ("align_tuned", C.c_uint32),
    ]


class NdtTuningEntry(C.Structure):
    _fields_ = [
        ("scan_bits", C.c_int32),
        ("batch_bits", C.c_int32),
        ("workgroups_per_alignment", C.c_int32),
        ("kernel_ms", C.c_float),
    ]


class NdtFilterParams(C.Structure):
    _fields_ = [
        ("min_distance", C.c_float),
        ("max_distance", C.c_float),
        ("min_z", C.c_float),
        ("max_z", C.c_float),
        ("downsample_resolution", C.c_float),
        ("downsample", C.c_int32),
    ]


class NdtFilterCounts(C.Structure):
    _fields_ = [
        ("kept", C.c_uint64),
        ("removed_by_distance", C.c_uint64),
        ("removed_by_z", C.c_uint64),
        ("removed_by_downsampling", C.c_uint64),
    ]


class NdtPhaseProfile(C.Structure):
    _fields_ = [
        ("derivative_us", C.c_float),
        ("exchange_us", C.c_float),
        ("solve_us", C.c_float),
        ("line_search_us", C.c_float),
        ("nvtl_us", C.c_float),
        ("total_us", C.c_float),
        ("iterations", C.c_uint32),
        ("line_search_evaluations", C.c_uint32),
        ("slowest_derivative_us", C.c_float),
        ("solve_setup_us", C.c_float),
        ("solve_step_us", C.c_float),
        ("eigen_solves", C.c_uint32),
        ("jacobi_sweeps", C.c_uint32),
    ]


# Every symbol declared in include/ndt_hip.h.
EXPORTS = [
    "ndt_hip_config_default", "ndt_hip_create", "ndt_hip_destroy", "ndt_hip_last_error",
    "ndt_hip_set_target", "ndt_hip_set_target_device", "ndt_hip_has_target", "ndt_hip_num_voxels",
    "ndt_hip_get_voxels", "ndt_hip_get_build_debug", "ndt_hip_lookup", "ndt_hip_derivatives",
    "ndt_hip_align", "ndt_hip_align_device", "ndt_hip_align_device_noalloc", "ndt_hip_align_batch", "ndt_hip_align_batch_device",
    "ndt_hip_align_scans", "ndt_hip_evaluate_nvtl", "ndt_hip_evaluate_transform_probability",
    "ndt_hip_per_point_scores", "ndt_hip_evaluate_nvtl_batch", "ndt_hip_set_regularization_pose",
    "ndt_hip_clear_regularization_pose", "ndt_hip_laplace_covariance_xy", "ndt_hip_laplace_covariance_6x6",
    "ndt_hip_get_timing", "ndt_hip_get_phase_profile", "ndt_hip_get_tuning",
    "ndt_hip_server_start", "ndt_hip_server_submit", "ndt_hip_server_poll", "ndt_hip_server_stop",
    "ndt_hip_server_set_idle_timeout", "ndt_hip_mlp_create", "ndt_hip_mlp_destroy", "ndt_hip_mlp_last_error", "ndt_hip_mlp_num_params",
    "ndt_hip_mlp_get_params", "ndt_hip_mlp_set_params", "ndt_hip_mlp_forward_device", "ndt_hip_mlp_forward_backward",
    "ndt_hip_mlp_train_step_device", "ndt_hip_bev_device", "ndt_hip_filter_params_default", "ndt_hip_filter_points",
    "ndt_hip_filter_points_device", "ndt_hip_voxel_downsample",
]

_lib = None


def load():
    """Loads libndt_hip.so (raises OSError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C cuda_ndt_matcher_amd/csrc")
    # PyTorch ships its own libamdhip64.so.7.  Loading torch first makes this
    # library bind to that same HIP runtime (one runtime per process, so
    # device pointers and streams from torch are valid here); loading this
    # library first would pull in a second copy from /opt/rocm under the
    # same soname, which torch then picks up instead of its own.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    sz = C.c_size_t
    rp = C.POINTER(NdtResult)
    lib.ndt_hip_config_default.argtypes = [C.POINTER(NdtConfig)]
    lib.ndt_hip_config_default.restype = None
    lib.ndt_hip_create.argtypes = [C.POINTER(NdtConfig), C.POINTER(vp)]
    lib.ndt_hip_destroy.argtypes = [vp]
    lib.ndt_hip_destroy.restype = None
    lib.ndt_hip_last_error.argtypes = [vp]
    lib.ndt_hip_last_error.restype = C.c_char_p
    lib.ndt_hip_set_target.argtypes = [vp, f32p, sz]
    lib.ndt_hip_set_target_device.argtypes = [vp, vp, sz]
    lib.ndt_hip_has_target.argtypes = [vp]
    lib.ndt_hip_num_voxels.argtypes = [vp, C.POINTER(sz)]
    lib.ndt_hip_get_voxels.argtypes = [vp, i32p, f32p, f32p, f32p, f32p, u32p]
    lib.ndt_hip_get_build_debug.argtypes = [vp, i32p, u32p, u64p]
    lib.ndt_hip_lookup.argtypes = [vp, C.c_int32, f32p, sz, C.c_uint32, u32p, i32p]
    lib.ndt_hip_derivatives.argtypes = [vp, f32p, sz, f64p, C.c_int32, C.POINTER(C.c_double), f64p, f64p,
                                        C.POINTER(C.c_uint32)]
    lib.ndt_hip_align.argtypes = [vp, f32p, sz, f64p, rp]
    lib.ndt_hip_align_device.argtypes = [vp, vp, sz, f64p, rp]
    lib.ndt_hip_align_device_noalloc.argtypes = [vp, vp, sz, f64p, vp]
    lib.ndt_hip_align_batch.argtypes = [vp, f32p, sz, f64p, sz, rp]
    lib.ndt_hip_align_batch_device.argtypes = [vp, vp, sz, f64p, sz, rp]
    lib.ndt_hip_align_scans.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(sz), f64p, sz, rp]
    for name in ("ndt_hip_evaluate_nvtl", "ndt_hip_evaluate_transform_probability"):
        getattr(lib, name).argtypes = [vp, f32p, sz, f64p, C.POINTER(C.c_double)]
    lib.ndt_hip_per_point_scores.argtypes = [vp, f32p, sz, f64p, f64p]
    lib.ndt_hip_evaluate_nvtl_batch.argtypes = [vp, f32p, sz, f64p, sz, f64p]
    lib.ndt_hip_set_regularization_pose.argtypes = [vp, f64p]
    lib.ndt_hip_clear_regularization_pose.argtypes = [vp]
    lib.ndt_hip_laplace_covariance_xy.argtypes = [f64p, f64p]
    lib.ndt_hip_laplace_covariance_6x6.argtypes = [f64p, f64p]
    lib.ndt_hip_get_timing.argtypes = [vp, C.POINTER(NdtTiming)]
    lib.ndt_hip_get_tuning.argtypes = [vp, C.POINTER(NdtTuningEntry), sz, C.POINTER(sz)]
    lib.ndt_hip_get_phase_profile.argtypes = [vp, C.POINTER(NdtPhaseProfile)]
    lib.ndt_hip_filter_params_default.argtypes = [C.POINTER(NdtFilterParams)]
    lib.ndt_hip_filter_params_default.restype = None
    lib.ndt_hip_filter_points.argtypes = [vp, f32p, sz, C.POINTER(NdtFilterParams), f32p, sz, C.POINTER(sz),
                                          C.POINTER(NdtFilterCounts)]
    lib.ndt_hip_filter_points_device.argtypes = [vp, vp, sz, C.POINTER(NdtFilterParams), vp, sz, C.POINTER(sz),
                                                 C.POINTER(NdtFilterCounts)]
    lib.ndt_hip_voxel_downsample.argtypes = [vp, f32p, sz, C.c_float, f32p, sz, C.POINTER(sz)]
    lib.ndt_hip_server_start.argtypes = [vp, C.c_int32, C.c_int32, sz, C.POINTER(vp)]
    lib.ndt_hip_server_submit.argtypes = [vp, f32p, sz, f64p, C.POINTER(C.c_uint64)]
    lib.ndt_hip_server_poll.argtypes = [vp, rp, C.POINTER(C.c_uint64), sz, C.POINTER(sz)]
    lib.ndt_hip_server_stop.argtypes = [vp]
    lib.ndt_hip_server_set_idle_timeout.argtypes = [vp, C.c_double]
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lib.ndt_hip_mlp_create.argtypes = [i32p, C.c_int32, C.c_uint64, C.POINTER(vp)]
    lib.ndt_hip_mlp_destroy.argtypes = [vp]
    lib.ndt_hip_mlp_destroy.restype = None
    lib.ndt_hip_mlp_last_error.argtypes = [vp]
    lib.ndt_hip_mlp_last_error.restype = C.c_char_p
    lib.ndt_hip_mlp_num_params.argtypes = [vp]
    lib.ndt_hip_mlp_num_params.restype = sz
    lib.ndt_hip_mlp_get_params.argtypes = [vp, f32p]
    lib.ndt_hip_mlp_set_params.argtypes = [vp, f32p]
    lib.ndt_hip_mlp_forward_device.argtypes = [vp, vp, sz, f32p]
    lib.ndt_hip_mlp_forward_backward.argtypes = [vp, f32p, sz, f32p, f32p, f32p]
    lib.ndt_hip_mlp_train_step_device.argtypes = [vp, vp, f32p, sz, C.c_float, C.POINTER(C.c_float)]
    lib.ndt_hip_bev_device.argtypes = [vp, vp, vp, sz, vp, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                       vp]
    _lib = lib
    return lib


def missing_exports() -> list:
    """Names declared in include/ndt_hip.h that the library does not export."""
    lib = load()
    out = []
    for name in EXPORTS:
        try:
            getattr(lib, name)
        except AttributeError:
            out.append(name)
    return out
