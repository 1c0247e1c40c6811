"""NdtScanMatcher: the reference crate's public API over libndt_hip.so.

Mirrors `ndt_cuda::NdtScanMatcher` (src/ndt_cuda/src/ndt.rs:41-1320): the
builder, set_target, align, evaluate_nvtl, evaluate_transform_probability,
compute_per_point_scores, evaluate_nvtl_batch, align_batch(_gpu),
align_parallel_scans and the regularisation pose, with the same argument
meaning and error classes.  Every call goes to the HIP library.
"""
from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .isometry import Isometry3

SEARCH_KDTREE, SEARCH_DIRECT26, SEARCH_DIRECT7, SEARCH_DIRECT1 = 0, 1, 2, 3
METRIC_POINT_TO_DISTRIBUTION, METRIC_POINT_TO_PLANE = 0, 1
PRECISION_FP64, PRECISION_FP32 = 0, 1

STATUS_NAMES = {0: "Converged", 1: "MaxIterations", 2: "NoCorrespondences", 3: "SingularHessian"}


class NdtError(RuntimeError):
    """An error returned by the library (code is the ndt_status value)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_lib.ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class AlignResult:
    """AlignResult (ndt.rs:220-248) plus the convergence status.  `pose` (an
    Isometry3) is formed from `pose_vector` when first read."""

    __slots__ = ("pose_vector", "converged", "status", "score", "transform_probability", "nvtl", "iterations",
                 "hessian", "num_correspondences", "oscillation_count", "pair_evaluations", "_pose")

    def __init__(self, pose_vector, converged, status, score, transform_probability, nvtl, iterations, hessian,
                 num_correspondences, oscillation_count, pose=None, pair_evaluations=0):
        self.pose_vector = pose_vector
        self.converged = converged
        self.status = status
        self.score = score
        self.transform_probability = transform_probability
        self.nvtl = nvtl
        self.iterations = iterations
        self.hessian = hessian
        self.num_correspondences = num_correspondences
        self.oscillation_count = oscillation_count
        self.pair_evaluations = pair_evaluations  # pairs over the Hessian passes (work done; not in the reference)
        self._pose = pose

    @property
    def pose(self) -> Isometry3:
        if self._pose is None:
            self._pose = Isometry3.from_pose_vector(self.pose_vector)
        return self._pose

    def __repr__(self) -> str:
        return (f"AlignResult(pose_vector={self.pose_vector!r}, status={self.status!r}, "
                f"iterations={self.iterations}, score={self.score!r}, nvtl={self.nvtl!r})")


@dataclass
class NdtScanMatcherConfig:
    """NdtScanMatcherConfig (ndt.rs:41-107) with the reference defaults."""

    resolution: float = 2.0
    max_iterations: int = 30
    trans_epsilon: float = 0.01
    step_size: float = 0.1
    outlier_ratio: float = 0.55
    regularization: float = 0.001
    use_line_search: bool = False
    min_points_per_voxel: int = 6
    use_gpu: bool = True
    regularization_enabled: bool = False
    regularization_scale_factor: float = 0.01
    distance_metric: int = METRIC_POINT_TO_DISTRIBUTION
    search_mode: int = SEARCH_KDTREE
    precision: int = PRECISION_FP64
    workgroups_per_alignment: int = 0
    profile_phases: bool = False
    cell_tables: bool = True
    use_graph: bool = False
    autotune: bool = False

    def to_c(self) -> _lib.NdtConfig:
        c = _lib.NdtConfig()
        _lib.load().ndt_hip_config_default(C.byref(c))
        c.resolution = float(self.resolution)
        c.max_iterations = int(self.max_iterations)
        c.trans_epsilon = float(self.trans_epsilon)
        c.step_size = float(self.step_size)
        c.outlier_ratio = float(self.outlier_ratio)
        c.regularization = float(self.regularization)
        c.use_line_search = int(bool(self.use_line_search))
        c.min_points_per_voxel = int(self.min_points_per_voxel)
        c.regularization_enabled = int(bool(self.regularization_enabled))
        c.regularization_scale_factor = float(self.regularization_scale_factor)
        c.distance_metric = int(self.distance_metric)
        c.search_mode = int(self.search_mode)
        c.precision = int(self.precision)
        c.workgroups_per_alignment = int(self.workgroups_per_alignment)
        c.profile_phases = int(bool(self.profile_phases))
        c.cell_tables = int(bool(self.cell_tables))
        c.use_graph = int(bool(self.use_graph))
        c.autotune = int(bool(self.autotune))
        return c


class NdtScanMatcherBuilder:
    """NdtScanMatcherBuilder (ndt.rs:111-210)."""

    def __init__(self):
        self.config = NdtScanMatcherConfig()

    def _set(self, **kw):
        for k, v in kw.items():
            setattr(self.config, k, v)
        return self

    def resolution(self, v: float):
        return self._set(resolution=v)

    def max_iterations(self, v: int):
        return self._set(max_iterations=v)

    def transformation_epsilon(self, v: float):
        return self._set(trans_epsilon=v)

    def step_size(self, v: float):
        return self._set(step_size=v)

    def outlier_ratio(self, v: float):
        return self._set(outlier_ratio=v)

    def regularization(self, v: float):
        return self._set(regularization=v)

    def use_line_search(self, v: bool):
        return self._set(use_line_search=v)

    def min_points_per_voxel(self, v: int):
        return self._set(min_points_per_voxel=v)

    def use_gpu(self, v: bool):
        return self._set(use_gpu=v)

    def regularization_enabled(self, v: bool):
        return self._set(regularization_enabled=v)

    def regularization_scale_factor(self, v: float):
        return self._set(regularization_scale_factor=v)

    def distance_metric(self, v: int):
        return self._set(distance_metric=v)

    def search_mode(self, v: int):
        return self._set(search_mode=v)

    def precision(self, v: int):
        return self._set(precision=v)

    def workgroups_per_alignment(self, v: int):
        return self._set(workgroups_per_alignment=v)

    def profile_phases(self, v: bool):
        return self._set(profile_phases=v)

    def cell_tables(self, v: bool):
        """Per-cell neighbour/slot tables for the lookups (identical results;
        hash + bitmap only when off)."""
        return self._set(cell_tables=v)

    def use_graph(self, v: bool):
        """Replay a HIP graph per launch shape for align* (same results)."""
        return self._set(use_graph=v)

    def autotune(self, v: bool):
        """Pick workgroups per alignment by timing candidates once per
        (scan size, batch size) bucket (with workgroups_per_alignment 0)."""
        return self._set(autotune=v)

    def build(self) -> "NdtScanMatcher":
        return NdtScanMatcher.with_config(self.config)


def _points(points) -> np.ndarray:
    a = np.asarray(points, dtype=np.float32)
    if a.size == 0:
        return np.zeros((0, 3), np.float32)
    return np.ascontiguousarray(a.reshape(-1, 3))


def _pose_vec(pose) -> np.ndarray:
    if isinstance(pose, Isometry3):
        return np.ascontiguousarray(pose.to_pose_vector(), dtype=np.float64)
    return np.ascontiguousarray(np.asarray(pose, dtype=np.float64).reshape(6))


_RESULT_DTYPE = None


def _results(arr, m: int) -> List[AlignResult]:
    """AlignResults of the first m entries of a ctypes ndt_result array, read
    through one structured numpy view (its dtype formed once: forming it from
    the ctypes type costs about 0.1 ms a call)."""
    global _RESULT_DTYPE
    if _RESULT_DTYPE is None:
        _RESULT_DTYPE = np.ctypeslib.as_array((_lib.NdtResult * 1)()).dtype
    a = np.frombuffer(arr, dtype=_RESULT_DTYPE, count=m)
    poses = np.array(a["pose"], dtype=np.float64)
    hess = np.array(a["hessian"], dtype=np.float64).reshape(m, 6, 6)
    status, conv = a["status"].tolist(), a["converged"].tolist()
    score, tp, nvtl = a["score"].tolist(), a["transform_probability"].tolist(), a["nvtl"].tolist()
    its, corr, osc = a["iterations"].tolist(), a["num_correspondences"].tolist(), a["oscillation_count"].tolist()
    pe = a["pair_evaluations"].tolist()
    return [AlignResult(poses[i], bool(conv[i]), STATUS_NAMES.get(status[i], str(status[i])), score[i], tp[i], nvtl[i],
                        its[i], hess[i], corr[i], osc[i], pair_evaluations=pe[i]) for i in range(m)]


def _result(r: _lib.NdtResult) -> AlignResult:
    """The AlignResult of one ndt_result, field by field (a few µs)."""
    return AlignResult(np.array(r.pose, dtype=np.float64), bool(r.converged), STATUS_NAMES.get(r.status, str(r.status)),
                       r.score, r.transform_probability, r.nvtl, r.iterations,
                       np.array(r.hessian, dtype=np.float64).reshape(6, 6), r.num_correspondences, r.oscillation_count,
                       pair_evaluations=r.pair_evaluations)


class NdtScanMatcher:
    """NdtScanMatcher (ndt.rs:255-1320) backed by the MI355X library.

    Poses may be given as Isometry3 (converted exactly as the reference does)
    or directly as the internal 6-vector [tx, ty, tz, roll, pitch, yaw].
    """

    def __init__(self, config: NdtScanMatcherConfig):
        self._lib = _lib.load()
        self._config = config
        h = C.c_void_p()
        cfg = config.to_c()
        rc = self._lib.ndt_hip_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise NdtError(rc, "ndt_hip_create failed (invalid or unsupported configuration, or no HIP device)")
        self._h = h

    @staticmethod
    def new(resolution: float) -> "NdtScanMatcher":
        return NdtScanMatcher(NdtScanMatcherConfig(resolution=resolution))

    @staticmethod
    def with_config(config: NdtScanMatcherConfig) -> "NdtScanMatcher":
        return NdtScanMatcher(config)

    @staticmethod
    def builder() -> NdtScanMatcherBuilder:
        return NdtScanMatcherBuilder()

    def config(self) -> NdtScanMatcherConfig:
        return self._config

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ndt_hip_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # At interpreter shutdown the HIP runtime (and a profiler's hooks) may
        # already be torn down: the handle is left to the process exit then.
        # Callers release handles with close() (or a with block) before that.
        # (sys itself may be cleared by then, hence the bound default.)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- internals ---------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._lib.ndt_hip_last_error(self._h)
            raise NdtError(rc, msg.decode() if msg else "")

    # -- target ------------------------------------------------------------
    def set_target(self, points) -> None:
        """set_target (ndt.rs:361-415): builds the voxel grid on the GPU."""
        p = _points(points)
        self._check(self._lib.ndt_hip_set_target(self._h, p, p.shape[0]))

    def set_target_device(self, ptr: int, n: int) -> None:
        self._check(self._lib.ndt_hip_set_target_device(self._h, C.c_void_p(ptr), n))

    def has_target(self) -> bool:
        return bool(self._lib.ndt_hip_has_target(self._h))

    def num_voxels(self) -> int:
        n = C.c_size_t()
        self._check(self._lib.ndt_hip_num_voxels(self._h, C.byref(n)))
        return int(n.value)

    def target_grid(self) -> Optional[dict]:
        """The voxel grid (Morton order): coords, means, covariances, inverses, axes, counts."""
        if not self.has_target():
            return None
        v = self.num_voxels()
        out = dict(coords=np.zeros((v, 3), np.int32), means=np.zeros((v, 3), np.float32),
                   covs=np.zeros((v, 3, 3), np.float32), invs=np.zeros((v, 3, 3), np.float32),
                   axes=np.zeros((v, 3), np.float32), counts=np.zeros(v, np.uint32))
        self._check(self._lib.ndt_hip_get_voxels(self._h, out["coords"], out["means"], out["covs"], out["invs"],
                                                 out["axes"], out["counts"]))
        return out

    def build_debug(self, n_points: int) -> dict:
        keys = np.zeros((n_points, 3), np.int32)
        order = np.zeros(n_points, np.uint32)
        skeys = np.zeros(n_points, np.uint64)
        self._check(self._lib.ndt_hip_get_build_debug(self._h, keys, order, skeys))
        return dict(point_keys=keys, sorted_index=order, sorted_keys=skeys)

    def lookup(self, queries, mode: int = SEARCH_KDTREE, max_k: int = 32) -> Tuple[np.ndarray, np.ndarray]:
        q = _points(queries)
        n = q.shape[0]
        cnt = np.zeros(n, np.uint32)
        idx = np.zeros((n, max_k), np.int32)
        self._check(self._lib.ndt_hip_lookup(self._h, mode, q, n, max_k, cnt, idx))
        return cnt, idx

    def derivatives(self, points, pose, compute_hessian: bool = True):
        p = _points(points)
        s = C.c_double()
        g = np.zeros(6)
        H = np.zeros(36)
        c = C.c_uint32()
        self._check(self._lib.ndt_hip_derivatives(self._h, p, p.shape[0], _pose_vec(pose), int(compute_hessian),
                                                  C.byref(s), g, H, C.byref(c)))
        return s.value, g, H.reshape(6, 6), c.value

    # -- alignment -----------------------------------------------------------
    def align(self, source_points, initial_guess) -> AlignResult:
        """align (ndt.rs:421-476)."""
        p = _points(source_points)
        r = _lib.NdtResult()
        self._check(self._lib.ndt_hip_align(self._h, p, p.shape[0], _pose_vec(initial_guess), C.byref(r)))
        return _result(r)

    def align_device(self, ptr: int, n: int, initial_guess) -> AlignResult:
        r = _lib.NdtResult()
        self._check(self._lib.ndt_hip_align_device(self._h, C.c_void_p(ptr), n, _pose_vec(initial_guess),
                                                   C.byref(r)))
        return _result(r)

    def align_batch(self, source_points, initial_poses: Sequence) -> List[AlignResult]:
        """align_batch / align_batch_gpu (ndt.rs:903-1001): one launch for all poses."""
        p = _points(source_points)
        m = len(initial_poses)
        if m == 0:
            if p.shape[0] == 0:
                raise NdtError(-2, "Source point cloud is empty")
            return []
        poses = np.ascontiguousarray(np.stack([_pose_vec(x) for x in initial_poses]).reshape(-1))
        res = (_lib.NdtResult * m)()
        self._check(self._lib.ndt_hip_align_batch(self._h, p, p.shape[0], poses, m, res))
        return _results(res, m)

    align_batch_gpu = align_batch

    def align_batch_device(self, ptr: int, n: int, initial_poses: Sequence) -> List[AlignResult]:
        m = len(initial_poses)
        poses = np.ascontiguousarray(np.stack([_pose_vec(x) for x in initial_poses]).reshape(-1))
        res = (_lib.NdtResult * m)()
        self._check(self._lib.ndt_hip_align_batch_device(self._h, C.c_void_p(ptr), n, poses, m, res))
        return _results(res, m)

    def align_parallel_scans(self, scans: Sequence[Tuple[object, object]]) -> List[AlignResult]:
        """align_parallel_scans (ndt.rs:1224-1320): uneven scans, one launch."""
        m = len(scans)
        if m == 0:
            return []
        arrays = [_points(s) for s, _ in scans]
        ptrs = (C.c_void_p * m)(*[a.ctypes.data for a in arrays])
        ns = (C.c_size_t * m)(*[a.shape[0] for a in arrays])
        poses = np.ascontiguousarray(np.stack([_pose_vec(p) for _, p in scans]).reshape(-1))
        res = (_lib.NdtResult * m)()
        self._check(self._lib.ndt_hip_align_scans(self._h, ptrs, ns, poses, m, res))
        return _results(res, m)

    def server(self, slots: int, workgroups_per_slot: int, max_points: int) -> "AlignServer":
        """Continuous-batching server over this matcher's map (see AlignServer)."""
        return AlignServer(self, slots, workgroups_per_slot, max_points)

    # -- scoring ---------------------------------------------------------------
    def evaluate_nvtl(self, source_points, pose) -> float:
        """evaluate_nvtl (ndt.rs:740-762)."""
        p = _points(source_points)
        out = C.c_double()
        self._check(self._lib.ndt_hip_evaluate_nvtl(self._h, p, p.shape[0], _pose_vec(pose), C.byref(out)))
        return out.value

    def evaluate_transform_probability(self, source_points, pose) -> float:
        """evaluate_transform_probability (ndt.rs:768-797)."""
        p = _points(source_points)
        out = C.c_double()
        self._check(self._lib.ndt_hip_evaluate_transform_probability(self._h, p, p.shape[0], _pose_vec(pose),
                                                                     C.byref(out)))
        return out.value

    def compute_per_point_scores(self, source_points, pose) -> np.ndarray:
        """compute_per_point_scores (ndt.rs:800-817)."""
        p = _points(source_points)
        out = np.zeros(p.shape[0])
        self._check(self._lib.ndt_hip_per_point_scores(self._h, p, p.shape[0], _pose_vec(pose), out))
        return out

    def evaluate_nvtl_batch(self, source_points, poses: Sequence) -> np.ndarray:
        """evaluate_nvtl_batch (ndt.rs:833-890)."""
        p = _points(source_points)
        m = len(poses)
        out = np.zeros(m)
        if m == 0:
            return out
        pv = np.ascontiguousarray(np.stack([_pose_vec(x) for x in poses]).reshape(-1))
        self._check(self._lib.ndt_hip_evaluate_nvtl_batch(self._h, p, p.shape[0], pv, m, out))
        return out

    # -- regularisation ------------------------------------------------------
    def set_regularization_pose(self, pose) -> None:
        self._check(self._lib.ndt_hip_set_regularization_pose(self._h, _pose_vec(pose)))

    def clear_regularization_pose(self) -> None:
        self._check(self._lib.ndt_hip_clear_regularization_pose(self._h))

    def phase_profile(self) -> dict:
        """Per-phase device time of the last alignment (profile_phases must be on)."""
        t = _lib.NdtPhaseProfile()
        self._check(self._lib.ndt_hip_get_phase_profile(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def timing(self) -> dict:
        t = _lib.NdtTiming()
        self._check(self._lib.ndt_hip_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def tuning(self) -> List[dict]:
        """The autotuner's table: one entry per (scan_bits, batch_bits) bucket."""
        cnt = C.c_size_t()
        self._check(self._lib.ndt_hip_get_tuning(self._h, None, 0, C.byref(cnt)))
        arr = (_lib.NdtTuningEntry * max(1, cnt.value))()
        self._check(self._lib.ndt_hip_get_tuning(self._h, arr, cnt.value, C.byref(cnt)))
        return [{k: getattr(arr[i], k) for k, _ in arr[i]._fields_} for i in range(cnt.value)]


def laplace_covariance_xy(hessian) -> np.ndarray:
    """estimate_xy_covariance_by_laplace (cuda_ndt_matcher/src/covariance.rs:82-98)."""
    out = np.zeros(4)
    rc = _lib.load().ndt_hip_laplace_covariance_xy(np.ascontiguousarray(hessian, np.float64).reshape(36), out)
    if rc != 0:
        raise NdtError(rc, "laplace")
    return out.reshape(2, 2)


def laplace_covariance_6x6(hessian) -> np.ndarray:
    out = np.zeros(36)
    rc = _lib.load().ndt_hip_laplace_covariance_6x6(np.ascontiguousarray(hessian, np.float64).reshape(36), out)
    if rc != 0:
        raise NdtError(rc, "laplace")
    return out.reshape(6, 6)


class AlignServer:
    """Continuous batching of alignments in one persistent launch.

    ``slots`` requests run at once, each on ``workgroups_per_slot`` cooperating
    workgroups; a request submitted while others run is posted to the first
    slot that frees up.  ``submit`` returns a ticket; ``poll`` returns the
    finished ``(ticket, AlignResult)`` pairs and admits queued requests.  A
    result equals ``align`` of the same scan and pose with
    ``workgroups_per_alignment = workgroups_per_slot``.  Use as a context
    manager (``stop`` runs on exit).  While it runs, the matcher's other
    calls raise NdtError: the server holds every co-resident workgroup.
    """

    def __init__(self, matcher: "NdtScanMatcher", slots: int, workgroups_per_slot: int, max_points: int):
        self._m = matcher
        self._lib = matcher._lib
        self._s = C.c_void_p()
        self._cap = int(slots)
        matcher._check(self._lib.ndt_hip_server_start(matcher._h, int(slots), int(workgroups_per_slot),
                                                      int(max_points), C.byref(self._s)))
        self._res = (_lib.NdtResult * self._cap)()
        self._tk = (C.c_uint64 * self._cap)()

    def submit(self, source_points, initial_guess) -> int:
        pts = _points(source_points)
        t = C.c_uint64()
        self._m._check(self._lib.ndt_hip_server_submit(self._s, pts, pts.shape[0], _pose_vec(initial_guess),
                                                       C.byref(t)))
        return int(t.value)

    def poll(self) -> List[Tuple[int, AlignResult]]:
        cnt = C.c_size_t()
        self._m._check(self._lib.ndt_hip_server_poll(self._s, self._res, self._tk, self._cap, C.byref(cnt)))
        return list(zip([int(t) for t in self._tk[:cnt.value]], _results(self._res, cnt.value)))

    def set_idle_timeout(self, seconds: float) -> None:
        """The launch ends when no submit or poll comes for this long (10 s by
        default); submit and poll then raise NdtError (HIP)."""
        self._m._check(self._lib.ndt_hip_server_set_idle_timeout(self._s, float(seconds)))

    def stop(self) -> None:
        if self._s:
            s, self._s = self._s, C.c_void_p()
            self._m._check(self._lib.ndt_hip_server_stop(s))

    def __enter__(self) -> "AlignServer":
        return self

    def __exit__(self, *exc) -> None:
        self.stop()
