"""Non-blocking map updates with two matchers.

Mirrors cuda_ndt_matcher/src/dual_ndt_manager.rs:1-357 (DualNdtManager): one
matcher serves alignments while the other rebuilds its voxel grid on a
background thread; swap_if_ready() exchanges them once the rebuild is done.
A rebuild requested while one runs is queued (the latest request wins) and
starts when the running one is swapped in.

Differences from the reference, both about resources and error reporting,
not results:
  * the two library handles are reused in turn instead of creating a fresh
    matcher per update (its device buffers only grow), and the rebuilt
    matcher's regularisation pose is cleared first, as a fresh matcher has
    none;
  * a failed rebuild is reported by the swap_if_ready() that collects it
    (the reference leaves the update flagged as in progress for good).

The library allows this: each handle has its own stream and no call
synchronises the whole device, so a build on one handle overlaps alignments
on the other.
"""
from __future__ import annotations

import threading
import time
from contextlib import contextmanager
from dataclasses import dataclass, replace
from typing import Callable, Optional

import numpy as np

from .matcher import NdtScanMatcher, NdtScanMatcherConfig


@dataclass
class UpdateStatus:
    """UpdateStatus (dual_ndt_manager.rs:31-41)."""
    in_progress: bool = False
    pending_points: int = 0
    swap_count: int = 0
    last_update_ms: float = 0.0


class DualMatcher:
    """DualNdtManager over two NdtScanMatcher handles."""

    def __init__(self, config: NdtScanMatcherConfig,
                 factory: Optional[Callable[[NdtScanMatcherConfig], NdtScanMatcher]] = None):
        make = factory or NdtScanMatcher
        self._config = config
        self._active = make(config)
        self._standby = make(config)
        self._lock = threading.RLock()   # the active matcher (the reference's RwLock)
        self._state = threading.Lock()   # the fields below
        self._thread: Optional[threading.Thread] = None
        self._ready = None
        self._error: Optional[BaseException] = None
        self._in_progress = False
        self._pending: Optional[np.ndarray] = None
        self._status = UpdateStatus()

    # -- background updates ------------------------------------------------
    def is_update_in_progress(self) -> bool:
        with self._state:
            return self._in_progress

    def get_status(self) -> UpdateStatus:
        with self._state:
            return replace(self._status)

    def start_background_update(self, points) -> bool:
        """start_background_update (dual_ndt_manager.rs:112-173): True if the
        rebuild started, False if one is running (the points are queued)."""
        pts = np.array(points, dtype=np.float32, copy=True).reshape(-1, 3)
        with self._state:
            if self._in_progress:
                self._pending = pts
                return False
            self._in_progress = True
            self._status.in_progress = True
            self._status.pending_points = int(pts.shape[0])
            standby = self._standby

            def run():
                t0 = time.perf_counter()
                try:
                    standby.clear_regularization_pose()
                    standby.set_target(pts)
                except BaseException as e:  # reported by swap_if_ready
                    with self._state:
                        self._error = e
                    return
                ms = (time.perf_counter() - t0) * 1e3
                with self._state:
                    self._status.last_update_ms = ms
                    self._ready = standby

            self._thread = threading.Thread(target=run, name="ndt-map-update", daemon=True)
            self._thread.start()
        return True

    def swap_if_ready(self) -> bool:
        """swap_if_ready (dual_ndt_manager.rs:182-240): True if a finished
        rebuild was swapped in.  A queued update then starts."""
        with self._state:
            t = self._thread
            if t is None or t.is_alive():
                return False
        t.join()
        with self._lock:
            with self._state:
                self._thread = None
                ready, err = self._ready, self._error
                self._ready = self._error = None
                if ready is not None:
                    self._active, self._standby = ready, self._active
                    self._status.swap_count += 1
                self._in_progress = False
                self._status.in_progress = False
                self._status.pending_points = 0
                pending, self._pending = self._pending, None
        if pending is not None:
            self.start_background_update(pending)
        if err is not None:
            raise err
        return ready is not None

    def wait_for_update(self, timeout: Optional[float] = None) -> bool:
        """Blocks until the running rebuild (if any) ends, then swap_if_ready()."""
        with self._state:
            t = self._thread
        if t is not None:
            t.join(timeout)
        return self.swap_if_ready()

    # -- the active matcher ----------------------------------------------------
    def set_target(self, points) -> None:
        """Blocking set_target on the active matcher (dual_ndt_manager.rs:246-249)."""
        with self._lock:
            self._active.set_target(points)

    def has_target(self) -> bool:
        with self._lock:
            return self._active.has_target()

    def align(self, source_points, initial_pose):
        """align (dual_ndt_manager.rs:259-271): swaps in a finished rebuild first."""
        self.swap_if_ready()
        with self._lock:
            return self._active.align(source_points, initial_pose)

    def evaluate_nvtl(self, source_points, pose) -> float:
        with self._lock:
            return self._active.evaluate_nvtl(source_points, pose)

    def evaluate_transform_probability(self, source_points, pose) -> float:
        with self._lock:
            return self._active.evaluate_transform_probability(source_points, pose)

    def evaluate_nvtl_batch(self, source_points, poses):
        with self._lock:
            return self._active.evaluate_nvtl_batch(source_points, poses)

    def align_batch(self, source_points, initial_poses):
        with self._lock:
            return self._active.align_batch(source_points, initial_poses)

    def set_regularization_pose(self, pose) -> None:
        with self._lock:
            self._active.set_regularization_pose(pose)

    def clear_regularization_pose(self) -> None:
        with self._lock:
            self._active.clear_regularization_pose()

    def is_regularization_enabled(self) -> bool:
        return bool(self._config.regularization_enabled)

    @contextmanager
    def lock(self):
        """lock (dual_ndt_manager.rs:349-356): the active matcher, after swap_if_ready()."""
        self.swap_if_ready()
        with self._lock:
            yield self._active

    def close(self) -> None:
        with self._state:
            t = self._thread
        if t is not None:
            t.join()
        for m in (self._active, self._standby):
            close = getattr(m, "close", None)
            if close:
                close()
