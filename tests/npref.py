"""Brute-force numpy checks used to pin the oracle (test helpers only).

Each helper recomputes one quantity from its definition with numpy, in the
simplest possible way (all voxels, no spatial index), so that a mistake in
the oracle's faster formulation shows up as a mismatch.
"""
import math

import numpy as np


def keys_f32(points, res):
    p = np.asarray(points, np.float32)
    return np.floor(p / np.float32(res)).astype(np.int64)


def kiddo_dist2(q, means):
    """((dx*dx) + dy*dy) + dz*dz in float32, every operation rounded."""
    q = np.asarray(q, np.float32)
    m = np.asarray(means, np.float32)
    d = (q[None, :] - m).astype(np.float32)
    s = (d[:, 0] * d[:, 0]).astype(np.float32)
    s = (s + (d[:, 1] * d[:, 1]).astype(np.float32)).astype(np.float32)
    s = (s + (d[:, 2] * d[:, 2]).astype(np.float32)).astype(np.float32)
    return s


def radius_neighbors(q, means, res):
    r = np.float32(res)
    d = kiddo_dist2(q, means)
    idx = np.nonzero(d < r * r)[0]
    return idx[np.argsort(d[idx], kind="stable")]


def quat_from_euler(roll, pitch, yaw):
    sr, cr = math.sin(roll * 0.5), math.cos(roll * 0.5)
    sp, cp = math.sin(pitch * 0.5), math.cos(pitch * 0.5)
    sy, cy = math.sin(yaw * 0.5), math.cos(yaw * 0.5)
    return (cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
            cr * cp * sy - sr * sp * cy)


def rot_zyx(roll, pitch, yaw):
    """Rz(yaw) Ry(pitch) Rx(roll) as an explicit matrix product."""
    cr, sr, cp, sp, cy, sy = math.cos(roll), math.sin(roll), math.cos(pitch), math.sin(pitch), math.cos(yaw), math.sin(yaw)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def rot_xyz(roll, pitch, yaw):
    """Rx(roll) Ry(pitch) Rz(yaw): the optimiser's transform."""
    cr, sr, cp, sp, cy, sy = math.cos(roll), math.sin(roll), math.cos(pitch), math.sin(pitch), math.cos(yaw), math.sin(yaw)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def gauss(res, outlier):
    c1 = 10.0 * (1.0 - outlier)
    c2 = outlier / res ** 3
    d3 = -math.log(c2)
    d1 = -math.log(c1 + c2) - d3
    d2 = -2.0 * math.log((-math.log(c1 * math.exp(-0.5) + c2) - d3) / d1)
    return d1, d2


def brute_scores(points, pose, grid_export, res, outlier):
    """Per-point neighbour scores with the Isometry transform, all voxels scanned."""
    d1, d2 = gauss(res, outlier)
    R = rot_zyx(pose[3], pose[4], pose[5])
    x = np.asarray(points, np.float64) @ R.T + np.asarray(pose[:3])
    means = grid_export["means"]
    invs = grid_export["invs"].astype(np.float64)
    out = []
    for i in range(x.shape[0]):
        nb = radius_neighbors(x[i].astype(np.float32), means, res)
        s = []
        for v in nb:
            d = x[i] - means[v].astype(np.float64)
            m = d @ (invs[v] @ d)
            s.append(-d1 * math.exp(-d2 * m / 2.0))
        out.append(s)
    return out


def half_cubic(length=20.0, interval=0.2):
    """make_half_cubic_pcd (src/ndt_cuda/src/test_utils.rs:19-47)."""
    n = int(np.float32(length) / np.float32(interval)) + 1
    pts = []
    for i in range(n):
        for j in range(n):
            u = np.float32(interval) * np.float32(j)
            v = np.float32(interval) * np.float32(i)
            pts += [[u, v, 0.0], [0.0, u, v], [u, 0.0, v]]
    return np.array(pts, np.float32)
