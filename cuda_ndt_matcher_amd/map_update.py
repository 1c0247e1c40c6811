"""Map tiles around the vehicle and when to rebuild the target.

Mirrors cuda_ndt_matcher/src/map_module.rs: MapUpdateModule (:81-361) keeps
the loaded tiles by id and combines them into the target cloud when the
vehicle has moved more than update_distance (2-D) since the last update or
the tiles changed; apply_differential is DynamicMapLoader::handle_response
(:525-583), which adds the tiles a map server sends and removes the ids it
lists (the request/response transport itself is the node's).  The combined
cloud feeds DualMatcher.start_background_update.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class DynamicMapParams:
    """DynamicMapParams (params.rs): metres."""
    update_distance: float = 20.0
    map_radius: float = 150.0
    lidar_radius: float = 100.0


@dataclass
class MapTile:
    """MapTile (map_module.rs:41-48)."""
    id: str
    center: Tuple[float, float, float]
    points: np.ndarray


@dataclass
class MapUpdateResult:
    """MapUpdateResult (map_module.rs:52-63)."""
    updated: bool
    tiles_loaded: int
    total_points: int
    distance_from_last_update: float
    update_time_ms: float


@dataclass
class MapStats:
    """MapStats (map_module.rs:67-79)."""
    tiles_loaded: int = 0
    total_points_full: int = 0
    total_points_filtered: int = 0
    last_update_position: Optional[Tuple[float, float, float]] = None
    needs_rebuild: bool = True


@dataclass
class MapLoaderStatus:
    """The counters of MapLoaderStatus (map_module.rs:376-391) that the
    differential update sets."""
    last_request_success: bool = False
    tiles_added: int = 0
    tiles_removed: int = 0
    points_added: int = 0
    last_update_time: float = 0.0


def _xyz(p) -> Tuple[float, float, float]:
    if hasattr(p, "x"):
        return (float(p.x), float(p.y), float(getattr(p, "z", 0.0)))
    return (float(p[0]), float(p[1]), float(p[2]) if len(p) > 2 else 0.0)


def _dist2d(a, b) -> float:
    return math.hypot(a[0] - b[0], a[1] - b[1])


def tile_from_points(tile_id: str, points) -> MapTile:
    """convert_cell_to_tile (map_module.rs:586-620): centre = f64 mean of the points."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    c = tuple(float(v) for v in p.astype(np.float64).mean(axis=0)) if p.shape[0] else (0.0, 0.0, 0.0)
    return MapTile(tile_id, c, p)


class MapUpdateModule:
    """MapUpdateModule (map_module.rs:81-361)."""

    def __init__(self, params: DynamicMapParams = DynamicMapParams()):
        self.params = params
        self._lock = threading.RLock()
        self._tiles: Dict[str, MapTile] = {}
        self._last: Optional[Tuple[float, float, float]] = None
        self._cached = np.zeros((0, 3), np.float32)
        self._needs_rebuild = True

    def should_update(self, position) -> bool:
        """No tiles, no previous update, or moved more than update_distance (:112-127)."""
        pos = _xyz(position)
        with self._lock:
            if not self._tiles or self._last is None:
                return True
            return _dist2d(pos, self._last) > self.params.update_distance

    def distance_from_last_update(self, position) -> float:
        with self._lock:
            return math.inf if self._last is None else _dist2d(_xyz(position), self._last)

    def out_of_map_range(self, position) -> bool:
        """distance + lidar_radius > map_radius (:142-152)."""
        with self._lock:
            if self._last is None:
                return True
            return _dist2d(_xyz(position), self._last) + self.params.lidar_radius > self.params.map_radius

    def add_tile(self, tile: MapTile) -> None:
        with self._lock:
            self._tiles[tile.id] = tile
            self._needs_rebuild = True

    def remove_tile(self, tile_id: str) -> bool:
        with self._lock:
            if self._tiles.pop(tile_id, None) is None:
                return False
            self._needs_rebuild = True
            return True

    def get_loaded_tile_ids(self) -> List[str]:
        with self._lock:
            return list(self._tiles.keys())

    def tile_count(self) -> int:
        with self._lock:
            return len(self._tiles)

    def update_map(self, position) -> MapUpdateResult:
        """update_map (:192-259): all loaded tiles' points combined (no
        distance filter: the map server selects tiles by radius)."""
        pos = _xyz(position)
        distance = self.distance_from_last_update(pos)
        with self._lock:
            if not self._needs_rebuild and not self.should_update(pos):
                return MapUpdateResult(False, len(self._tiles), int(self._cached.shape[0]), distance, 0.0)
            t0 = time.perf_counter()
            parts = [t.points for t in self._tiles.values()]
            self._cached = (np.concatenate(parts).astype(np.float32, copy=False) if parts
                            else np.zeros((0, 3), np.float32))
            self._last = pos
            self._needs_rebuild = False
            ms = (time.perf_counter() - t0) * 1e3
            return MapUpdateResult(True, len(self._tiles), int(self._cached.shape[0]), distance, ms)

    def get_stats(self) -> MapStats:
        with self._lock:
            return MapStats(len(self._tiles), int(sum(t.points.shape[0] for t in self._tiles.values())),
                            int(self._cached.shape[0]), self._last, self._needs_rebuild)

    def check_and_update(self, position) -> Optional[np.ndarray]:
        """The combined points if update_map rebuilt them (:284-291)."""
        return self.get_map_points() if self.update_map(position).updated else None

    def get_map_points(self) -> Optional[np.ndarray]:
        with self._lock:
            return None if self._cached.shape[0] == 0 else self._cached.copy()

    def clear(self) -> None:
        with self._lock:
            self._tiles.clear()
            self._cached = np.zeros((0, 3), np.float32)
            self._last = None
            self._needs_rebuild = True

    def load_full_map(self, points) -> None:
        """One tile 'full_map' replacing everything (:322-356)."""
        self.clear()
        self.add_tile(tile_from_points("full_map", points))

    def apply_differential(self, new_tiles: Iterable[Tuple[str, object]], ids_to_remove: Sequence[str],
                           status: Optional[MapLoaderStatus] = None) -> MapLoaderStatus:
        """handle_response (:525-583): add (id, points) tiles, then remove ids.
        `points` may be an (n, 3) array or a PointCloud2."""
        from .pointcloud import from_pointcloud2

        st = status or MapLoaderStatus()
        added = removed = pts = 0
        for tile_id, cloud in new_tiles:
            p = from_pointcloud2(cloud) if hasattr(cloud, "point_step") else cloud
            tile = tile_from_points(tile_id, p)
            pts += tile.points.shape[0]
            self.add_tile(tile)
            added += 1
        for tile_id in ids_to_remove:
            removed += self.remove_tile(tile_id)
        st.last_request_success = True
        st.tiles_added, st.tiles_removed, st.points_added = added, removed, pts
        st.last_update_time = time.time()
        return st
