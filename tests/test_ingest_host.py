"""Either side of the hot path, on the CPU: the sensor point filter's oracle
against the reference's own cases (filtering/cpu.rs and pointcloud.rs tests)
and a numpy brute force; PointCloud2 conversion; the map tile module
(map_module.rs tests); the dual matcher's update/swap logic with stub
matchers."""
import math
import threading
import time

import numpy as np
import pytest

from cuda_ndt_matcher_amd import pointcloud as PC
from cuda_ndt_matcher_amd.dual import DualMatcher
from cuda_ndt_matcher_amd.map_update import DynamicMapParams, MapTile, MapUpdateModule
from oracle import oracle as O

FLT_MAX = O.FLT_MAX

# (points, params, expected points (None: not checked) in order, counts
# [kept, distance, z, downsampling]); from filtering/cpu.rs:105-275 and
# pointcloud.rs:348-417.  Downsampled expectations are sorted by cell.
KAT_CASES = [
    ([], {}, [], [0, 0, 0, 0]),
    ([[0.5, 0, 0], [2, 0, 0], [5, 0, 0]], dict(min_distance=1.0), [[2, 0, 0], [5, 0, 0]], [2, 1, 0, 0]),
    ([[5, 0, 0], [50, 0, 0], [150, 0, 0]], dict(max_distance=100.0), [[5, 0, 0], [50, 0, 0]], [2, 1, 0, 0]),
    ([[1, 0, -5], [1, 0, 0], [1, 0, 5], [1, 0, 15]], dict(min_z=-2.0, max_z=10.0), [[1, 0, 0], [1, 0, 5]],
     [2, 0, 2, 0]),
    ([[0.5, 0, 0], [5, 0, -5], [5, 0, 0], [5, 0, 15], [150, 0, 0]],
     dict(min_distance=1.0, max_distance=100.0, min_z=-2.0, max_z=10.0), [[5, 0, 0]], [1, 2, 2, 0]),
    ([[1, 0, 0], [5, 0, 0], [10, 0, 0], [15, 0, 0]], dict(min_distance=3.0, max_distance=12.0),
     [[5, 0, 0], [10, 0, 0]], [2, 2, 0, 0]),
    ([[1, 0, -2], [1, 0, 0.5], [1, 0, 1.5], [1, 0, 5]], dict(min_z=-1.0, max_z=3.0), [[1, 0, 0.5], [1, 0, 1.5]],
     [2, 0, 2, 0]),
    ([[0.1, 0, 0], [0.2, 0, 0], [5, 0, 0]], dict(downsample_resolution=1.0), None, [3, 0, 0, 1]),
    ([[1, 2, 3], [4, 5, 6]], {}, [[1, 2, 3], [4, 5, 6]], [2, 0, 0, 0]),
    ([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.3, 0.3, 0.3], [0.4, 0.4, 0.4]], dict(downsample_resolution=1.0), None,
     [4, 0, 0, 3]),
    ([[0.1, 0.1, 0.1], [1.1, 0.1, 0.1], [0.1, 1.1, 0.1], [0.1, 0.1, 1.1]], dict(downsample_resolution=1.0), None,
     [4, 0, 0, 0]),
    ([[-0.1, -0.1, -0.1], [-0.2, -0.2, -0.2]], dict(downsample_resolution=1.0), None, [2, 0, 0, 1]),
    ([[1, 2, 3]], dict(downsample_resolution=0.0), [[1, 2, 3]], [1, 0, 0, 0]),
    ([[1, 2, 3]], dict(downsample_resolution=-1.0), [[1, 2, 3]], [1, 0, 0, 0]),
    # a NaN or infinite resolution is not <= 0, so the cells are floor(p * NaN)
    # or floor(p * 0) = 0 after the saturating cast: one cell for everything
    ([[1, 2, 3], [-7, 5, 0.5]], dict(downsample_resolution=float("nan")), [[-3, 3.5, 1.75]], [2, 0, 0, 1]),
    ([[1, 2, 3], [-7, 5, 0.5]], dict(downsample_resolution=float("inf")), [[-3, 3.5, 1.75]], [2, 0, 0, 1]),
]


def brute_filter(points, min_distance=0.0, max_distance=FLT_MAX, min_z=-FLT_MAX, max_z=FLT_MAX,
                 downsample_resolution=None):
    """numpy restatement: vectorised tests, cells grouped with np.unique and
    summed in input order with np.add.at (unbuffered, sequential)."""
    p = np.asarray(points, np.float32).reshape(-1, 3)
    f32 = np.float32
    with np.errstate(over="ignore", invalid="ignore"):
        d = (p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1]) + p[:, 2] * p[:, 2]
        far = (d < f32(min_distance) * f32(min_distance)) | (d > f32(max_distance) * f32(max_distance))
        badz = ~far & ((p[:, 2] < f32(min_z)) | (p[:, 2] > f32(max_z)))
    kept = p[~far & ~badz]
    counts = [kept.shape[0], int(far.sum()), int(badz.sum()), 0]
    if downsample_resolution is None or not downsample_resolution > 0 or kept.shape[0] == 0:
        return kept, counts
    inv = f32(1.0) / f32(downsample_resolution)
    cells = np.floor(kept * inv).astype(np.int64)
    uniq, inverse = np.unique(cells, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    sums = np.zeros((uniq.shape[0], 3))
    np.add.at(sums, inverse, kept.astype(np.float64))
    cnt = np.bincount(inverse, minlength=uniq.shape[0]).astype(np.float64)
    out = (sums / cnt[:, None]).astype(np.float32)
    counts[3] = kept.shape[0] - uniq.shape[0]
    return out, counts  # lexicographic cell order, like the oracle


@pytest.mark.parametrize("k", range(len(KAT_CASES)))
def test_oracle_filter_reference_cases(k):
    pts, kw, want, counts = KAT_CASES[k]
    got, c = O.filter_points(pts, **kw)
    assert c == counts
    if want is not None:
        np.testing.assert_array_equal(got, np.asarray(want, np.float32).reshape(-1, 3))


def test_oracle_downsample_known_centroids():
    # cpu.rs:211-260: one voxel -> (0.25, 0.25, 0.25); negative coords -> -0.15
    got, _ = O.filter_points([[0.1] * 3, [0.2] * 3, [0.3] * 3, [0.4] * 3], downsample_resolution=1.0)
    np.testing.assert_allclose(got, [[0.25, 0.25, 0.25]], atol=1e-5)
    got, _ = O.filter_points([[-0.1] * 3, [-0.2] * 3], downsample_resolution=1.0)
    np.testing.assert_allclose(got, [[-0.15, -0.15, -0.15]], atol=1e-5)
    got, c = O.filter_points([[0.1, 0, 0], [0.2, 0, 0], [5, 0, 0]], downsample_resolution=1.0)
    np.testing.assert_allclose(got, [[0.15, 0, 0], [5, 0, 0]], atol=1e-6)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_filter_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 20, (20000, 3)).astype(np.float32)
    p[::97] = p[::97].round(0)  # points on cell faces
    kw = dict(min_distance=2.0, max_distance=45.0, min_z=-10.0, max_z=15.0, downsample_resolution=0.7)
    got, c = O.filter_points(p, **kw)
    want, wc = brute_filter(p, **kw)
    assert c == wc
    np.testing.assert_array_equal(got, want)
    kw.pop("downsample_resolution")
    got, c = O.filter_points(p, **kw)
    want, wc = brute_filter(p, **kw)
    assert c == wc
    np.testing.assert_array_equal(got, want)


def test_oracle_filter_nan_passes_the_tests():
    # NaN compares false, so the CPU path keeps the point (cpu.rs:19-33)
    got, c = O.filter_points([[np.nan, 0, 0], [1, 0, 0]], min_distance=0.5, max_distance=10.0, min_z=-1.0, max_z=1.0)
    assert c == [2, 0, 0, 0] and np.isnan(got[0, 0])


# -- PointCloud2 ---------------------------------------------------------------

def test_pointcloud2_reference_cases():
    # pointcloud.rs:331-345
    msg = PC.to_pointcloud2([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    assert (msg.point_step, msg.row_step, msg.width, msg.height, msg.is_bigendian) == (12, 24, 2, 1, False)
    np.testing.assert_array_equal(PC.from_pointcloud2(msg), [[1, 2, 3], [4, 5, 6]])
    assert PC.from_pointcloud2(PC.to_pointcloud2(np.zeros((0, 3)))).shape == (0, 3)


def test_pointcloud2_layouts_and_errors():
    rng = np.random.default_rng(3)
    p = rng.normal(size=(50, 3)).astype(np.float32)
    # 32-byte points, fields out of order, with padding and an intensity field
    rec = np.zeros(50, dtype=np.dtype({"names": ["i", "z", "x", "y"], "formats": ["<f4"] * 4,
                                       "offsets": [0, 8, 16, 24], "itemsize": 32}))
    rec["x"], rec["y"], rec["z"], rec["i"] = p[:, 0], p[:, 1], p[:, 2], 7.0
    p[5, 1] = np.nan
    rec["y"][5] = np.nan
    rec["z"][9] = np.inf
    fields = [PC.PointField("intensity", 0), PC.PointField("z", 8), PC.PointField("x", 16), PC.PointField("y", 24)]
    msg = PC.PointCloud2(5, 10, fields, False, 32, 320, rec.tobytes())
    got = PC.from_pointcloud2(msg)
    keep = np.ones(50, bool)
    keep[[5, 9]] = False
    np.testing.assert_array_equal(got, p[keep])
    with pytest.raises(ValueError, match="Missing 'z'"):
        PC.from_pointcloud2(PC.PointCloud2(1, 1, fields[2:], False, 32, 32, bytes(32)))
    with pytest.raises(ValueError, match="too short"):
        PC.from_pointcloud2(PC.PointCloud2(1, 51, fields, False, 32, 1632, rec.tobytes()))
    # rgb writer: bits of 0x00RRGGBB as the fourth float, missing colours 0
    m = PC.to_pointcloud2_with_rgb(p[:3], [0x00FF8040, 0x00010203])
    raw = np.frombuffer(m.data, dtype="<u4").reshape(3, 4)
    assert list(raw[:, 3]) == [0x00FF8040, 0x00010203, 0]
    np.testing.assert_array_equal(PC.from_pointcloud2(m), p[:3])


# -- map tiles (map_module.rs:638-836) -------------------------------------------

def _tile(pts, tid="tile_1"):
    return MapTile(tid, (0.0, 0.0, 0.0), np.asarray(pts, np.float32).reshape(-1, 3))


def test_map_module_reference_cases():
    P = (0.0, 0.0, 0.0)
    m = MapUpdateModule(DynamicMapParams())
    assert m.should_update(P)
    m.add_tile(_tile([[0, 0, 0]]))
    m.update_map(P)
    assert not m.should_update((10.0, 0.0, 0.0))
    assert m.should_update((25.0, 0.0, 0.0))
    assert not m.out_of_map_range((40.0, 0.0, 0.0))
    assert m.out_of_map_range((60.0, 0.0, 0.0))
    e = MapUpdateModule(DynamicMapParams())
    e.update_map(P)
    assert e.should_update((5.0, 0.0, 0.0))  # no tiles: always update
    a = MapUpdateModule()
    a.add_tile(_tile([[0, 0, 0], [1, 1, 1]]))
    assert a.tile_count() == 1 and "tile_1" in a.get_loaded_tile_ids()
    assert a.remove_tile("tile_1") and a.tile_count() == 0 and not a.remove_tile("tile_1")
    r = MapUpdateModule(DynamicMapParams(map_radius=10.0))
    r.add_tile(_tile([[0, 0, 0], [5, 0, 0], [15, 0, 0], [0, 20, 0]]))
    res = r.update_map(P)
    assert res.updated and res.total_points == 4 and res.update_time_ms >= 0.0
    c = MapUpdateModule(DynamicMapParams(update_distance=5.0))
    c.add_tile(_tile([[0, 0, 0], [1, 1, 1]]))
    assert c.check_and_update(P) is not None
    assert c.check_and_update(P) is None
    assert c.check_and_update((10.0, 0.0, 0.0)) is not None
    s = MapUpdateModule()
    s.add_tile(_tile([[0, 0, 0], [1, 1, 1], [2, 2, 2]]))
    st = s.get_stats()
    assert (st.tiles_loaded, st.total_points_full, st.total_points_filtered, st.needs_rebuild) == (1, 3, 0, True)
    s.update_map(P)
    st = s.get_stats()
    assert not st.needs_rebuild and st.total_points_filtered > 0
    f = MapUpdateModule()
    f.add_tile(_tile([[9, 9, 9]], "other"))
    f.load_full_map([[0, 0, 0], [2, 4, 6]])
    assert f.get_loaded_tile_ids() == ["full_map"]


def test_map_module_differential_update():
    m = MapUpdateModule()
    a = np.arange(12, dtype=np.float32).reshape(4, 3)
    b = -a
    st = m.apply_differential([("a", a), ("b", PC.to_pointcloud2(b))], [])
    assert (st.tiles_added, st.tiles_removed, st.points_added) == (2, 0, 8)
    np.testing.assert_array_equal(m.check_and_update((0.0, 0.0, 0.0)), np.concatenate([a, b]))
    st = m.apply_differential([("c", a[:1])], ["a", "missing"])
    assert (st.tiles_added, st.tiles_removed, st.points_added) == (1, 1, 1)
    np.testing.assert_array_equal(m.check_and_update((0.0, 0.0, 0.0)), np.concatenate([b, a[:1]]))


# -- dual matcher (dual_ndt_manager.rs) with stub matchers ---------------------------

class _StubMatcher:
    def __init__(self, config):
        self.map = None
        self.reg = "set"
        self.gate = None
        self.fail = False

    def set_target(self, pts):
        if self.gate is not None:
            assert self.gate.wait(10)
        if self.fail:
            raise RuntimeError("build failed")
        self.map = np.asarray(pts).copy()

    def clear_regularization_pose(self):
        self.reg = None

    def has_target(self):
        return self.map is not None

    def align(self, scan, pose):
        return ("aligned", None if self.map is None else float(self.map.sum()))


def test_dual_matcher_swap_and_queue():
    d = DualMatcher(None, factory=_StubMatcher)
    d.set_target(np.ones((2, 3)))
    gate = threading.Event()
    d._standby.gate = gate
    assert d.start_background_update(np.full((3, 3), 2.0))
    assert d.is_update_in_progress() and d.get_status().pending_points == 3
    assert d.align(None, None) == ("aligned", 6.0)  # the old map while rebuilding
    assert not d.start_background_update(np.full((1, 3), 5.0))  # queued
    assert not d.swap_if_ready()
    gate.set()
    d._thread.join(10)
    # the next align swaps, and the queued update starts on the other handle
    assert d.align(None, None) == ("aligned", 18.0)
    assert d.get_status().swap_count == 1 and d.is_update_in_progress()
    assert d.wait_for_update(10)
    assert d.align(None, None) == ("aligned", 15.0)
    st = d.get_status()
    assert (st.swap_count, st.in_progress, st.pending_points) == (2, False, 0) and st.last_update_ms >= 0.0
    with d.lock() as m:
        assert m.reg is None  # a rebuilt matcher starts without a regularisation pose


def test_dual_matcher_reports_a_failed_rebuild():
    d = DualMatcher(None, factory=_StubMatcher)
    d.set_target(np.ones((1, 3)))
    d._standby.fail = True
    assert d.start_background_update(np.zeros((1, 3)))
    d._thread.join(10)
    with pytest.raises(RuntimeError, match="build failed"):
        d.swap_if_ready()
    assert not d.is_update_in_progress()
    assert d.align(None, None) == ("aligned", 3.0)
