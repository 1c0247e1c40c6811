"""sensor_msgs/PointCloud2 <-> AoS float32 points.

Mirrors cuda_ndt_matcher/src/pointcloud.rs:97-285: from_pointcloud2 reads the
'x', 'y', 'z' fields as little-endian float32 at their offsets, point i at
i * point_step (height * width points), and drops points with a non-finite
coordinate; to_pointcloud2 / to_pointcloud2_with_rgb write a dense
little-endian cloud of 12 / 16 bytes per point.  The message classes are plain
data holders with the ROS message's field names, so a ROS message object with
those attributes can be passed as is.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

FLOAT32 = 7  # sensor_msgs/PointField datatype


@dataclass
class PointField:
    name: str
    offset: int
    datatype: int = FLOAT32
    count: int = 1


@dataclass
class PointCloud2:
    height: int
    width: int
    fields: List[PointField]
    is_bigendian: bool
    point_step: int
    row_step: int
    data: bytes
    is_dense: bool = True
    header: Optional[Any] = None


def from_pointcloud2(msg) -> np.ndarray:
    """from_pointcloud2 (pointcloud.rs:134-166): (k, 3) float32."""
    data = bytes(msg.data) if not isinstance(msg.data, (bytes, bytearray, memoryview)) else msg.data
    if len(data) == 0:
        return np.zeros((0, 3), np.float32)
    off = {}
    for f in msg.fields:  # the last field of a name wins, as in the reference's loop
        if f.name in ("x", "y", "z"):
            off[f.name] = int(f.offset)
    for a in ("x", "y", "z"):
        if a not in off:
            raise ValueError(f"Missing '{a}' field")
    n = int(msg.width) * int(msg.height)
    step = int(msg.point_step)
    if len(data) < n * step:
        raise ValueError(f"PointCloud2 data too short: {len(data)} < {n * step}")
    if n == 0:
        return np.zeros((0, 3), np.float32)
    dt = np.dtype({"names": ["x", "y", "z"], "formats": ["<f4"] * 3, "offsets": [off["x"], off["y"], off["z"]],
                   "itemsize": step})
    rec = np.frombuffer(data, dtype=dt, count=n)
    pts = np.stack([rec["x"], rec["y"], rec["z"]], axis=1).astype(np.float32)
    return np.ascontiguousarray(pts[np.isfinite(pts).all(axis=1)])


def _xyz_fields(extra: bool) -> List[PointField]:
    f = [PointField("x", 0), PointField("y", 4), PointField("z", 8)]
    if extra:
        f.append(PointField("rgb", 12))
    return f


def to_pointcloud2(points, header=None) -> PointCloud2:
    """to_pointcloud2 (pointcloud.rs:241-280)."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3)).astype("<f4")
    n = p.shape[0]
    return PointCloud2(1, n, _xyz_fields(False), False, 12, 12 * n, p.tobytes(), True, header)


def to_pointcloud2_with_rgb(points, rgb_values, header=None) -> PointCloud2:
    """to_pointcloud2_with_rgb (pointcloud.rs:183-238): rgb packed 0x00RRGGBB,
    its bits stored as the fourth float; missing colours are 0."""
    p = np.asarray(points, np.float32).reshape(-1, 3)
    n = p.shape[0]
    rgb = np.zeros(n, np.uint32)
    v = np.asarray(rgb_values, np.uint32).reshape(-1)[:n]
    rgb[:v.shape[0]] = v
    rec = np.zeros(n, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("rgb", "<u4")])
    rec["x"], rec["y"], rec["z"], rec["rgb"] = p[:, 0], p[:, 1], p[:, 2], rgb
    return PointCloud2(1, n, _xyz_fields(True), False, 16, 16 * n, rec.tobytes(), True, header)
