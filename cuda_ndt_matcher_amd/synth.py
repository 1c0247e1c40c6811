"""Seeded synthetic urban scenes, lidar scans and initial-guess perturbations.

Wraps libndt_synth.so (csrc/synth.cpp).  The workloads:

* small  : 200 x 200 m, 40 boxes, 100 poles, 200k map points; 32-beam lidar
           downsampled at 1.0 m (about 1.5k points); guess N(0, 0.5 m) on
           x, y, z and N(0, 2 deg) on yaw.
* full   : 1 x 1 km, boxes and poles scaled by area (1000 / 2500), 5M map
           points; 128-beam lidar, 1800 azimuth steps, downsampled at 0.3 m to
           about 29k points; guess N(0, 1.0 m) on x, y, z and N(0, 3 deg) on roll,
           pitch and yaw.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libndt_synth.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise OSError(f"{_LIB_PATH} is not built")
        lib = C.CDLL(_LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        lib.synth_scene_create.restype = C.c_void_p
        lib.synth_scene_create.argtypes = [C.c_uint64, C.c_double, C.c_uint32, C.c_uint32]
        lib.synth_scene_destroy.argtypes = [C.c_void_p]
        lib.synth_scene_sample.restype = C.c_size_t
        lib.synth_scene_sample.argtypes = [C.c_void_p, C.c_uint64, C.c_size_t, C.c_double, f32p]
        lib.synth_scene_lidar.restype = C.c_size_t
        lib.synth_scene_lidar.argtypes = [C.c_void_p, C.c_uint64, f64p, C.c_uint32, C.c_double, C.c_double,
                                          C.c_uint32, C.c_double, C.c_double, f32p]
        lib.synth_voxel_downsample.restype = C.c_size_t
        lib.synth_voxel_downsample.argtypes = [f32p, C.c_size_t, C.c_float, f32p]
        _lib = lib
    return _lib


def _splitmix(x: int) -> int:
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def rng(seed: int, stream: int) -> np.random.Generator:
    """numpy Generator keyed by (seed, stream) through SplitMix64."""
    return np.random.default_rng(_splitmix((seed * 0x9E3779B1 + stream) & ((1 << 64) - 1)))


class Scene:
    def __init__(self, seed: int, extent: float, n_boxes: int, n_poles: int):
        self._lib = _load()
        self.seed, self.extent = seed, extent
        self._h = self._lib.synth_scene_create(seed, extent, n_boxes, n_poles)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.synth_scene_destroy(h)
            self._h = None

    def sample(self, n: int, noise: float = 0.02, seed: int | None = None) -> np.ndarray:
        out = np.zeros((n, 3), np.float32)
        k = self._lib.synth_scene_sample(self._h, self.seed if seed is None else seed, n, noise, out)
        return out[:k]

    def lidar(self, x: float, y: float, z: float, yaw: float, beams: int, az_steps: int = 1800,
              vmin: float = -25.0, vmax: float = 15.0, max_range: float = 100.0, noise: float = 0.02,
              seed: int = 0) -> np.ndarray:
        out = np.zeros((beams * az_steps, 3), np.float32)
        pose = np.array([x, y, z, yaw], np.float64)
        k = self._lib.synth_scene_lidar(self._h, seed, pose, beams, vmin, vmax, az_steps, max_range, noise, out)
        return out[:k]


def voxel_downsample(points, leaf: float) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    out = np.zeros_like(p)
    k = _load().synth_voxel_downsample(p, p.shape[0], leaf, out)
    return out[:k]


@dataclass
class Workload:
    name: str
    map_points: np.ndarray
    scan: np.ndarray
    true_pose: np.ndarray   # [x, y, z, roll, pitch, yaw]
    guess: np.ndarray       # perturbed initial pose
    resolution: float = 2.0


WORKLOADS = {
    # name: (extent, boxes, poles, map points, beams, leaf, sigma_t, sigma_deg, rpy perturbed)
    "small": (200.0, 40, 100, 200_000, 32, 1.0, 0.5, 2.0, False),
    "full": (1000.0, 1000, 2500, 5_000_000, 128, 0.3, 1.0, 3.0, True),
}


def sensor_pose(seed: int, extent: float) -> np.ndarray:
    r = rng(seed, 11)
    half = extent * 0.5 * 0.6
    return np.array([r.uniform(-half, half), r.uniform(-half, half), 1.8, 0.0, 0.0, r.uniform(-math.pi, math.pi)])


def perturb(pose: np.ndarray, seed: int, sigma_t: float, sigma_deg: float, rpy: bool) -> np.ndarray:
    r = rng(seed, 12)
    g = pose.copy()
    g[:3] += r.normal(0.0, sigma_t, 3)
    d = np.deg2rad(sigma_deg)
    if rpy:
        g[3:] += r.normal(0.0, d, 3)
    else:
        g[5] += r.normal(0.0, d)
    return g


_SCENES = {}


def dense_scan(name: str, seed: int, leaf: float) -> np.ndarray:
    """The workload's lidar scan (same scene, sensor pose and beams)
    downsampled at a finer `leaf`, for requests larger than the workload's."""
    extent, nb, npl, _, beams, _, _, _, _ = WORKLOADS[name]
    wl = make_workload(name, seed)
    scene = Scene(seed, extent, nb, npl)
    raw = scene.lidar(wl.true_pose[0], wl.true_pose[1], wl.true_pose[2], wl.true_pose[5], beams, seed=seed + 1)
    return voxel_downsample(raw, leaf)


def raw_scan(name: str, seed: int) -> np.ndarray:
    """The workload's lidar scan before downsampling (sensor frame)."""
    extent, nb, npl, _, beams, _, _, _, _ = WORKLOADS[name]
    wl = make_workload(name, seed)
    scene = Scene(seed, extent, nb, npl)
    return scene.lidar(wl.true_pose[0], wl.true_pose[1], wl.true_pose[2], wl.true_pose[5], beams, seed=seed + 1)


def sweep_scans(seed: int, sizes, leaf: float = 0.1):
    """Scans of the given sizes over the small workload's map (the precision
    sweep): one 128-beam scan from the small workload's sensor pose,
    downsampled at `leaf`, then sub-sampled without replacement to each size
    (seeded).  Returns (workload, [scan per size])."""
    wl = make_workload("small", seed)
    extent = WORKLOADS["small"][0]
    scene = Scene(seed, extent, WORKLOADS["small"][1], WORKLOADS["small"][2])
    raw = scene.lidar(wl.true_pose[0], wl.true_pose[1], wl.true_pose[2], wl.true_pose[5], 128, seed=seed + 1)
    dense = voxel_downsample(raw, leaf)
    r = rng(seed, 13)
    out = []
    for n in sizes:
        if n > dense.shape[0]:
            raise ValueError(f"only {dense.shape[0]} points available for a {n}-point scan")
        idx = np.sort(r.choice(dense.shape[0], n, replace=False))
        out.append(np.ascontiguousarray(dense[idx]))
    return wl, out


def make_workload(name: str, seed: int = 0, map_points: int | None = None) -> Workload:
    extent, nb, npl, nmap, beams, leaf, st, sd, rpy = WORKLOADS[name]
    n = nmap if map_points is None else map_points
    key = (name, seed, n)
    if key in _SCENES:
        return _SCENES[key]
    scene = Scene(seed, extent, nb, npl)
    pts = scene.sample(n, 0.02, seed)
    pose = sensor_pose(seed + 1, extent)
    raw = scene.lidar(pose[0], pose[1], pose[2], pose[5], beams, seed=seed + 1)
    scan = voxel_downsample(raw, leaf)
    guess = perturb(pose, seed + 2, st, sd, rpy)
    wl = Workload(name, pts, scan, pose, guess)
    _SCENES[key] = wl
    return wl
