"""Sensor point filtering on the GPU.

Mirrors src/ndt_cuda/src/filtering/mod.rs (FilterParams :13-31, FilterResult
:34-41, GpuPointFilter::filter :61-230) and the node's
cuda_ndt_matcher/src/pointcloud.rs:59-95 (filter_sensor_points): points
outside [min_distance, max_distance] of the sensor origin or outside
[min_z, max_z] are dropped, then the rest are optionally reduced to one
centroid per cell of a voxel grid.

Every cloud runs through the library's kernels (the reference switches to its
CPU path below 10k points; here the GPU path has the CPU path's semantics:
a removed point is counted under the first test it fails, distance before z,
and the distance is x*x + y*y + z*z rounded after every operation).
Downsampled cells come out in the library's key order, not HashMap order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .matcher import NdtError, NdtScanMatcher, NdtScanMatcherConfig, _points

FLT_MAX = float(np.finfo(np.float32).max)


@dataclass
class FilterParams:
    """FilterParams / PointFilterParams (filtering/mod.rs:13-31,
    pointcloud.rs:9-32); downsample_resolution None means no downsampling."""
    min_distance: float = 0.0
    max_distance: float = FLT_MAX
    min_z: float = -FLT_MAX
    max_z: float = FLT_MAX
    downsample_resolution: Optional[float] = None

    def to_c(self) -> _lib.NdtFilterParams:
        p = _lib.NdtFilterParams()
        p.min_distance = self.min_distance
        p.max_distance = self.max_distance
        p.min_z = self.min_z
        p.max_z = self.max_z
        p.downsample = int(self.downsample_resolution is not None)
        p.downsample_resolution = float(self.downsample_resolution) if p.downsample else 0.0
        return p


PointFilterParams = FilterParams


@dataclass
class FilterResult:
    """FilterResult (filtering/mod.rs:34-41)."""
    points: Optional[np.ndarray]
    removed_by_distance: int
    removed_by_z: int
    removed_by_downsampling: int
    used_gpu: bool = True


def _result(points, c: _lib.NdtFilterCounts) -> FilterResult:
    return FilterResult(points, int(c.removed_by_distance), int(c.removed_by_z), int(c.removed_by_downsampling), True)


class GpuPointFilter:
    """GpuPointFilter (filtering/mod.rs:43-230) over the library's filter
    kernels.  Uses the given matcher's handle (and stream), or a handle of its
    own."""

    def __init__(self, matcher: Optional[NdtScanMatcher] = None):
        self._own = matcher is None
        self._m = matcher if matcher is not None else NdtScanMatcher(NdtScanMatcherConfig())

    @staticmethod
    def new() -> "GpuPointFilter":
        return GpuPointFilter()

    def close(self) -> None:
        if self._own and self._m is not None:
            self._m.close()
        self._m = None

    def filter(self, points, params: FilterParams) -> FilterResult:
        p = _points(points)
        n = p.shape[0]
        out = np.empty((max(n, 1), 3), np.float32)
        cnt = _lib.NdtFilterCounts()
        got = C.c_size_t()
        cp = params.to_c()
        self._m._check(self._m._lib.ndt_hip_filter_points(self._m._h, p, n, C.byref(cp), out, n, C.byref(got),
                                                          C.byref(cnt)))
        return _result(out[:got.value].copy(), cnt)

    def filter_device(self, ptr: int, n: int, params: FilterParams, out_ptr: int, out_cap: int):
        """Device memory in and out (3 floats per point): returns the number of
        points written and the counts (FilterResult with points None)."""
        cnt = _lib.NdtFilterCounts()
        got = C.c_size_t()
        cp = params.to_c()
        rc = self._m._lib.ndt_hip_filter_points_device(self._m._h, C.c_void_p(ptr), n, C.byref(cp),
                                                       C.c_void_p(out_ptr), out_cap, C.byref(got), C.byref(cnt))
        if rc != 0:
            msg = self._m._lib.ndt_hip_last_error(self._m._h)
            raise NdtError(rc, (msg.decode() if msg else "") + f" (needed {got.value})")
        return int(got.value), _result(None, cnt)

    def voxel_downsample(self, points, resolution: float) -> np.ndarray:
        """voxel_downsample_cpu (filtering/cpu.rs:67-100) on the GPU."""
        p = _points(points)
        n = p.shape[0]
        out = np.empty((max(n, 1), 3), np.float32)
        got = C.c_size_t()
        self._m._check(self._m._lib.ndt_hip_voxel_downsample(self._m._h, p, n, resolution, out, n, C.byref(got)))
        return out[:got.value].copy()


def filter_sensor_points(points, params: FilterParams, gpu_filter: Optional[GpuPointFilter] = None) -> FilterResult:
    """filter_sensor_points (pointcloud.rs:59-95)."""
    f = gpu_filter or GpuPointFilter()
    try:
        return f.filter(points, params)
    finally:
        if gpu_filter is None:
            f.close()
