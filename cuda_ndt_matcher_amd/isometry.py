"""Isometry3 and the pose-vector conversions the reference applies at its API.

The reference's align() takes an Isometry3<f64>, turns it into the internal
[tx, ty, tz, roll, pitch, yaw] vector with nalgebra's euler_angles
(isometry_to_pose_vector, src/ndt_cuda/src/optimization/types.rs:164-174) and
turns the result back with UnitQuaternion::from_euler_angles
(pose_vector_to_isometry, types.rs:156-161).  Those conversions use
R = Rz(yaw) Ry(pitch) Rx(roll) while the optimiser's own transform is
Rx Ry Rz (derivatives/cpu.rs:339-372); keeping both here reproduces the
reference exactly.
"""
from __future__ import annotations

import math

import numpy as np


def quat_from_euler(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """nalgebra UnitQuaternion::from_euler_angles as (w, i, j, k)."""
    sr, cr = math.sin(roll * 0.5), math.cos(roll * 0.5)
    sp, cp = math.sin(pitch * 0.5), math.cos(pitch * 0.5)
    sy, cy = math.sin(yaw * 0.5), math.cos(yaw * 0.5)
    return np.array([cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy,
                     cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy])


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    w, i, j, k = q
    ww, ii, jj, kk = w * w, i * i, j * j, k * k
    ij, wk, wj, ik, jk, wi = i * j * 2, w * k * 2, w * j * 2, i * k * 2, j * k * 2, w * i * 2
    return np.array([[ww + ii - jj - kk, ij - wk, wj + ik],
                     [wk + ij, ww - ii + jj - kk, jk - wi],
                     [ik - wj, wi + jk, ww - ii - jj + kk]])


def matrix_to_euler(R: np.ndarray):
    """Rotation3::euler_angles (roll, pitch, yaw) as nalgebra computes it."""
    r20 = R[2, 0]
    if abs(r20) < 1.0:
        pitch = -math.asin(r20)
        c = math.cos(pitch)
        roll = math.atan2(R[2, 1] / c, R[2, 2] / c)
        yaw = math.atan2(R[1, 0] / c, R[0, 0] / c)
    elif r20 <= -1.0:
        roll, pitch, yaw = math.atan2(R[0, 1], R[0, 2]), math.pi / 2, 0.0
    else:
        roll, pitch, yaw = -math.atan2(-R[0, 1], -R[0, 2]), -math.pi / 2, 0.0
    return roll, pitch, yaw


class Isometry3:
    """Rigid transform: rotation matrix R (3x3) and translation t (3,)."""

    __slots__ = ("R", "t")

    def __init__(self, R=None, t=None):
        self.R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64).reshape(3, 3)
        self.t = np.zeros(3) if t is None else np.asarray(t, dtype=np.float64).reshape(3)

    @staticmethod
    def identity() -> "Isometry3":
        return Isometry3()

    @staticmethod
    def translation(x: float, y: float, z: float) -> "Isometry3":
        return Isometry3(t=[x, y, z])

    @staticmethod
    def from_parts(t, roll: float, pitch: float, yaw: float) -> "Isometry3":
        return Isometry3(quat_to_matrix(quat_from_euler(roll, pitch, yaw)), t)

    @staticmethod
    def from_pose_vector(pose) -> "Isometry3":
        """pose_vector_to_isometry (optimization/types.rs:156-161)."""
        p = np.asarray(pose, dtype=np.float64).reshape(6)
        return Isometry3.from_parts(p[:3], p[3], p[4], p[5])

    def to_pose_vector(self) -> np.ndarray:
        """isometry_to_pose_vector (optimization/types.rs:164-174)."""
        roll, pitch, yaw = matrix_to_euler(self.R)
        return np.array([self.t[0], self.t[1], self.t[2], roll, pitch, yaw])

    def euler_angles(self):
        return matrix_to_euler(self.R)

    def transform_points(self, pts) -> np.ndarray:
        p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
        return p @ self.R.T + self.t

    def inverse(self) -> "Isometry3":
        return Isometry3(self.R.T, -self.R.T @ self.t)

    def __mul__(self, other: "Isometry3") -> "Isometry3":
        return Isometry3(self.R @ other.R, self.R @ other.t + self.t)

    def __repr__(self) -> str:
        return f"Isometry3(t={self.t.tolist()}, rpy={list(self.euler_angles())})"
