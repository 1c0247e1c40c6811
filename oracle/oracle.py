"""ctypes wrapper of the CPU oracle (libndt_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is a C++/FP64 restatement of the
reference crate's CPU path.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NDT_ORACLE_LIB selects another build of the same source (the
# AddressSanitizer/UBSan build that `make -C oracle sanitize` writes).
_LIB_PATH = os.environ.get("NDT_ORACLE_LIB") or os.path.join(_HERE, "libndt_oracle.so")

SEARCH_KDTREE, SEARCH_DIRECT26, SEARCH_DIRECT7, SEARCH_DIRECT1, SEARCH_CELL27 = 0, 1, 2, 3, 4


class Params(C.Structure):
    _fields_ = [
        ("resolution", C.c_double),
        ("outlier_ratio", C.c_double),
        ("max_iterations", C.c_uint32),
        ("trans_epsilon", C.c_double),
        ("step_size", C.c_double),
        ("use_line_search", C.c_int32),
        ("search_mode", C.c_int32),
        ("distance_metric", C.c_int32),
        ("regularization_enabled", C.c_int32),
        ("regularization_scale_factor", C.c_double),
        ("has_regularization_pose", C.c_int32),
        ("regularization_translation", C.c_double * 3),
        ("fp64_voxels", C.c_int32),
        ("noise", C.c_double),
        ("noise_seed", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("pose", C.c_double * 6),
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("score", C.c_double),
        ("transform_probability", C.c_double),
        ("nvtl", C.c_double),
        ("hessian", C.c_double * 36),
        ("num_correspondences", C.c_uint32),
        ("oscillation_count", C.c_uint32),
    ]


SCORE_FN = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


NATIVE_LIB_PATH = os.path.join(_HERE, "libndt_oracle_native.so")
NATIVE_FLAGS = "-O3 -march=native -ffp-contract=off"


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path=_LIB_PATH):
    if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(path)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    lib.oracle_gauss.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.oracle_point_keys.argtypes = [f32p, C.c_size_t, C.c_float, i32p]
    lib.oracle_grid_build.restype = C.c_void_p
    lib.oracle_grid_build.argtypes = [f32p, C.c_size_t, C.c_float, C.c_uint32, C.c_float]
    lib.oracle_grid_free.argtypes = [C.c_void_p]
    lib.oracle_grid_size.restype = C.c_size_t
    lib.oracle_grid_size.argtypes = [C.c_void_p]
    lib.oracle_grid_export.argtypes = [C.c_void_p, i32p, f32p, f32p, f32p, f32p, u32p, f64p, f64p]
    lib.oracle_search.argtypes = [C.c_void_p, C.c_int32, f32p, C.c_size_t, u32p, i32p, C.c_uint32]
    lib.oracle_derivatives.argtypes = [C.c_void_p, f32p, C.c_size_t, f64p, C.POINTER(Params), C.c_int32,
                                       C.POINTER(C.c_double), f64p, f64p, C.POINTER(C.c_uint32)]
    lib.oracle_newton_step.restype = C.c_int32
    lib.oracle_newton_step.argtypes = [f64p, f64p, C.c_double, C.c_double, f64p]
    lib.oracle_sym_eigen3.argtypes = [f64p, f64p, f64p]
    lib.oracle_sym_eigen6.argtypes = [f64p, f64p, f64p]
    lib.oracle_align.argtypes = [C.c_void_p, f32p, C.c_size_t, f64p, C.POINTER(Params), C.POINTER(Result)]
    for name in ("oracle_nvtl", "oracle_transform_probability", "oracle_nvtl_cell27"):
        fn = getattr(lib, name)
        fn.restype = C.c_double
        fn.argtypes = [C.c_void_p, f32p, C.c_size_t, f64p, C.POINTER(Params)]
    lib.oracle_per_point_scores.argtypes = [C.c_void_p, f32p, C.c_size_t, f64p, C.POINTER(Params), f64p]
    lib.oracle_more_thuente.argtypes = [SCORE_FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_uint32,
                                        C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_double)]
    lib.oracle_regularization.argtypes = [C.POINTER(Params), f64p, C.c_uint32, C.POINTER(C.c_double), f64p, f64p]
    lib.oracle_oscillation.restype = C.c_uint32
    lib.oracle_oscillation.argtypes = [f64p, C.c_size_t]
    lib.oracle_laplace_xy.argtypes = [f64p, f64p]
    lib.oracle_pose_to_matrix.argtypes = [f64p, f64p]
    lib.oracle_matrix_to_pose.argtypes = [f64p, f64p]
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    lib.oracle_filter_points.restype = C.c_size_t
    lib.oracle_filter_points.argtypes = [f32p, C.c_size_t, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                                         C.c_float, f32p, u64p]
    return lib


_lib = None
_native = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def native_lib():
    """The same restatement built for the host it runs on (-O3 -march=native,
    still without contraction, so its results equal the parity build's): the
    CPU-baseline timing build.  Compiled on first use (about 10 s); returns
    None when the compiler is unavailable."""
    global _native
    if _native is None:
        try:
            subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, capture_output=True)
            _native = _load(NATIVE_LIB_PATH)
        except (OSError, subprocess.CalledProcessError):
            return None
    return _native


def _pts(points) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
    return a


def _pose(pose) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(pose, dtype=np.float64).reshape(6))


def params(resolution=2.0, outlier_ratio=0.55, max_iterations=30, trans_epsilon=0.01, step_size=0.1,
           use_line_search=False, search_mode=SEARCH_KDTREE, distance_metric=0, regularization_enabled=False,
           regularization_scale_factor=0.01, regularization_translation=None, fp64_voxels=False, noise=0.0,
           noise_seed=0) -> Params:
    p = Params()
    p.resolution = resolution
    p.outlier_ratio = outlier_ratio
    p.max_iterations = max_iterations
    p.trans_epsilon = trans_epsilon
    p.step_size = step_size
    p.use_line_search = int(use_line_search)
    p.search_mode = search_mode
    p.distance_metric = distance_metric
    p.regularization_enabled = int(regularization_enabled)
    p.regularization_scale_factor = regularization_scale_factor
    if regularization_translation is not None:
        p.has_regularization_pose = 1
        for i in range(3):
            p.regularization_translation[i] = float(regularization_translation[i])
    p.fp64_voxels = int(fp64_voxels)
    p.noise = noise
    p.noise_seed = noise_seed
    return p


def gauss(resolution: float, outlier_ratio: float):
    d1, d2 = C.c_double(), C.c_double()
    lib().oracle_gauss(resolution, outlier_ratio, C.byref(d1), C.byref(d2))
    return d1.value, d2.value


def point_keys(points, resolution: float) -> np.ndarray:
    p = _pts(points)
    out = np.zeros((p.shape[0], 3), dtype=np.int32)
    lib().oracle_point_keys(p, p.shape[0], resolution, out)
    return out


@dataclass
class OracleResult:
    pose: np.ndarray
    status: int
    iterations: int
    score: float
    transform_probability: float
    nvtl: float
    hessian: np.ndarray
    num_correspondences: int
    oscillation_count: int

    @property
    def converged(self) -> bool:
        return self.status == 0


class Grid:
    """Voxel grid built by the oracle (voxels sorted by coordinate)."""

    def __init__(self, points, resolution: float = 2.0, min_points: int = 6, eigen_ratio: float = 0.01,
                 library=None):
        p = _pts(points)
        self.resolution = float(resolution)
        self._lib = library if library is not None else lib()
        self._h = self._lib.oracle_grid_build(p, p.shape[0], np.float32(resolution), min_points,
                                              np.float32(eigen_ratio))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.oracle_grid_free(h)
            self._h = None

    def __len__(self) -> int:
        return int(self._lib.oracle_grid_size(self._h))

    def export(self) -> dict:
        v = len(self)
        out = dict(coords=np.zeros((v, 3), np.int32), means=np.zeros((v, 3), np.float32),
                   covs=np.zeros((v, 3, 3), np.float32), invs=np.zeros((v, 3, 3), np.float32),
                   axes=np.zeros((v, 3), np.float32), counts=np.zeros(v, np.uint32),
                   means64=np.zeros((v, 3), np.float64), invs64=np.zeros((v, 3, 3), np.float64))
        self._lib.oracle_grid_export(self._h, out["coords"], out["means"], out["covs"], out["invs"], out["axes"],
                                 out["counts"], out["means64"], out["invs64"])
        return out

    def search(self, queries, mode: int = SEARCH_KDTREE, max_k: int = 32):
        q = _pts(queries)
        n = q.shape[0]
        cnt = np.zeros(n, np.uint32)
        idx = np.zeros((n, max_k), np.int32)
        self._lib.oracle_search(self._h, mode, q, n, cnt, idx, max_k)
        return cnt, idx

    def derivatives(self, points, pose, p: Params, compute_hessian: bool = True):
        pts = _pts(points)
        g = np.zeros(6)
        H = np.zeros(36)
        s = C.c_double()
        c = C.c_uint32()
        self._lib.oracle_derivatives(self._h, pts, pts.shape[0], _pose(pose), C.byref(p), int(compute_hessian),
                                 C.byref(s), g, H, C.byref(c))
        return s.value, g, H.reshape(6, 6), c.value

    def align(self, points, pose, p: Params) -> OracleResult:
        pts = _pts(points)
        r = Result()
        self._lib.oracle_align(self._h, pts, pts.shape[0], _pose(pose), C.byref(p), C.byref(r))
        return OracleResult(np.array(r.pose[:]), r.status, r.iterations, r.score, r.transform_probability, r.nvtl,
                            np.array(r.hessian[:]).reshape(6, 6), r.num_correspondences, r.oscillation_count)

    def nvtl(self, points, pose, p: Params) -> float:
        pts = _pts(points)
        return self._lib.oracle_nvtl(self._h, pts, pts.shape[0], _pose(pose), C.byref(p))

    def transform_probability(self, points, pose, p: Params) -> float:
        pts = _pts(points)
        return self._lib.oracle_transform_probability(self._h, pts, pts.shape[0], _pose(pose), C.byref(p))

    def nvtl_cell27(self, points, pose, p: Params) -> float:
        pts = _pts(points)
        return self._lib.oracle_nvtl_cell27(self._h, pts, pts.shape[0], _pose(pose), C.byref(p))

    def per_point_scores(self, points, pose, p: Params) -> np.ndarray:
        pts = _pts(points)
        out = np.zeros(pts.shape[0])
        self._lib.oracle_per_point_scores(self._h, pts, pts.shape[0], _pose(pose), C.byref(p), out)
        return out


def newton_step(grad, hess, min_eigenvalue=-1e-3, tolerance=1e-10):
    d = np.zeros(6)
    ok = lib().oracle_newton_step(np.ascontiguousarray(grad, np.float64), np.ascontiguousarray(hess, np.float64).ravel(),
                                  min_eigenvalue, tolerance, d)
    return d if ok else None


def sym_eigen(a):
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n))
    if n == 3:
        lib().oracle_sym_eigen3(a.ravel(), ev, V.reshape(-1))
    elif n == 6:
        lib().oracle_sym_eigen6(a.ravel(), ev, V.reshape(-1))
    else:
        raise ValueError("only 3x3 and 6x6")
    return ev, V


def more_thuente(fn, initial_score, initial_derivative, initial_step, step_max=0.1, step_min=1e-9, mu=1e-4, nu=0.9,
                 max_iterations=10):
    def cb(alpha, s, d, _user):
        a, b = fn(alpha)
        s[0] = a
        d[0] = b

    c_cb = SCORE_FN(cb)
    st, fv = C.c_double(), C.c_double()
    conv = C.c_int32()
    its = C.c_uint32()
    lib().oracle_more_thuente(c_cb, None, initial_score, initial_derivative, initial_step, step_max, step_min, mu,
                              nu, max_iterations, C.byref(st), C.byref(conv), C.byref(its), C.byref(fv))
    return dict(step_length=st.value, converged=bool(conv.value), iterations=its.value, final_value=fv.value)


def regularization(p: Params, pose, corr: int):
    s = C.c_double()
    g = np.zeros(6)
    H = np.zeros(36)
    lib().oracle_regularization(C.byref(p), _pose(pose), corr, C.byref(s), g, H)
    return s.value, g, H.reshape(6, 6)


def oscillation(translations) -> int:
    t = np.ascontiguousarray(np.asarray(translations, np.float64).reshape(-1, 3))
    return int(lib().oracle_oscillation(t.ravel(), t.shape[0]))


def laplace_xy(hess) -> np.ndarray:
    out = np.zeros(4)
    lib().oracle_laplace_xy(np.ascontiguousarray(hess, np.float64).ravel(), out)
    return out.reshape(2, 2)


def pose_to_matrix(pose) -> np.ndarray:
    m = np.zeros(12)
    lib().oracle_pose_to_matrix(_pose(pose), m)
    return m.reshape(3, 4)


def matrix_to_pose(m) -> np.ndarray:
    out = np.zeros(6)
    lib().oracle_matrix_to_pose(np.ascontiguousarray(m, np.float64).reshape(12), out)
    return out


FLT_MAX = float(np.finfo(np.float32).max)


def filter_points(points, min_distance=0.0, max_distance=FLT_MAX, min_z=-FLT_MAX, max_z=FLT_MAX,
                  downsample_resolution=None):
    """filter_points_cpu (filtering/cpu.rs:7-58) with voxel_downsample_cpu
    (:67-100) when downsample_resolution is not None.  Returns the points
    ((k, 3) float32; downsampled cells in lexicographic cell order) and the
    counts (kept before downsampling, removed by distance, by z, by
    downsampling)."""
    p = _pts(points)
    out = np.zeros((max(1, p.shape[0]), 3), np.float32)
    counts = np.zeros(4, np.uint64)
    ds = downsample_resolution is not None
    k = lib().oracle_filter_points(p.reshape(-1), p.shape[0], min_distance, max_distance, min_z, max_z, int(ds),
                                   float(downsample_resolution) if ds else 0.0, out.reshape(-1), counts)
    return out[:k].copy(), [int(c) for c in counts]
