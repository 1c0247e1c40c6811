"""Either side of the hot path on the GPU: the sensor point filter against the
oracle (bit-exact points and counts), its device form feeding an alignment,
its error paths, and the dual matcher rebuilding one map while the other
handle aligns."""
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O
from oracle import parity
from tests.test_ingest_host import KAT_CASES

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


def _rows(a):
    a = np.asarray(a, np.float32).reshape(-1, 3)
    return a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))]


def _params(kw):
    from cuda_ndt_matcher_amd.filtering import FilterParams

    return FilterParams(**kw)


@pytest.fixture(scope="module")
def gpu_filter():
    from cuda_ndt_matcher_amd.filtering import GpuPointFilter

    f = GpuPointFilter()
    yield f
    f.close()


def _check(gpu_filter, pts, kw):
    r = gpu_filter.filter(pts, _params(kw))
    want, wc = O.filter_points(pts, **kw)
    got_counts = [wc[0], r.removed_by_distance, r.removed_by_z, r.removed_by_downsampling]
    assert got_counts == wc
    if kw.get("downsample_resolution") is not None and kw["downsample_resolution"] > 0:
        np.testing.assert_array_equal(_rows(r.points), _rows(want))
    else:
        np.testing.assert_array_equal(r.points, want)  # input order
    return r


@pytest.mark.parametrize("k", range(len(KAT_CASES)))
def test_filter_reference_cases(gpu_filter, k):
    pts, kw, _, counts = KAT_CASES[k]
    _case(f"filter reference case {k}")
    r = _check(gpu_filter, pts, kw)
    assert r.used_gpu


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 100003])
def test_filter_random_clouds(gpu_filter, n):
    _case(f"filter random n={n}")
    rng = np.random.default_rng(n)
    p = rng.normal(0, 25, (n, 3)).astype(np.float32)
    p[::53] = p[::53].round(0)  # on cell faces
    kw = dict(min_distance=1.5, max_distance=60.0, min_z=-12.0, max_z=20.0)
    _check(gpu_filter, p, kw)
    kw["downsample_resolution"] = 0.5
    _check(gpu_filter, p, kw)


def test_filter_raw_lidar_scan(gpu_filter):
    """The 128-beam scan of the 5M-point workload (about 230k points): the
    node's preprocessing, then the same through voxel_downsample."""
    from cuda_ndt_matcher_amd import synth

    raw = synth.raw_scan("full", 0)
    _case(f"filter raw lidar scan {raw.shape[0]} pts")
    kw = dict(min_distance=1.0, max_distance=100.0, min_z=-3.0, max_z=40.0, downsample_resolution=0.3)
    r = _check(gpu_filter, raw, kw)
    assert 20000 < r.points.shape[0] < 60000
    want, _ = O.filter_points(raw, downsample_resolution=0.3)
    np.testing.assert_array_equal(_rows(gpu_filter.voxel_downsample(raw, 0.3)), _rows(want))


def test_filter_device_feeds_align():
    """filter_points_device output aligned in place equals the alignment of
    the host-filtered scan."""
    import torch

    from cuda_ndt_matcher_amd import NdtScanMatcher, synth
    from cuda_ndt_matcher_amd.filtering import FilterParams, GpuPointFilter

    wl = synth.make_workload("small", 0)
    raw = synth.raw_scan("small", 0)
    _case(f"filter device -> align, raw {raw.shape[0]} pts")
    m = NdtScanMatcher.builder().resolution(2.0).max_iterations(30).transformation_epsilon(0.01).build()
    m.set_target(wl.map_points)
    f = GpuPointFilter(m)
    prm = FilterParams(min_distance=1.0, max_distance=80.0, downsample_resolution=1.0)
    d_in = torch.from_numpy(raw).to("cuda:0")
    d_out = torch.empty_like(d_in)
    torch.cuda.synchronize()
    k, res = f.filter_device(d_in.data_ptr(), raw.shape[0], prm, d_out.data_ptr(), raw.shape[0])
    host = f.filter(raw, prm).points
    assert k == host.shape[0]
    np.testing.assert_array_equal(d_out[:k].cpu().numpy(), host)
    a = m.align_device(d_out.data_ptr(), k, wl.guess)
    b = m.align(host, wl.guess)
    assert np.array_equal(a.pose_vector, b.pose_vector) and a.iterations == b.iterations
    # too small an output: refused, with the count needed
    from cuda_ndt_matcher_amd import NdtError

    with pytest.raises(NdtError, match=f"needed {k}"):
        f.filter_device(d_in.data_ptr(), raw.shape[0], prm, d_out.data_ptr(), k - 1)
    # cells beyond 2^20 from the origin: refused, and the handle still works
    far = raw.copy()
    far[0] = (3e6, 0.0, 0.0)
    with pytest.raises(NdtError, match="2\\^20"):
        f.filter(far, FilterParams(downsample_resolution=1.0))
    c = m.align(host, wl.guess)
    assert np.array_equal(c.pose_vector, b.pose_vector)


def test_dual_matcher_rebuilds_while_aligning(full_wl, full_oracle, oracle_mod):
    """A 5M-point map is built on the standby handle while the active one
    keeps aligning on the 200k-point map.  Those alignments equal a separate
    matcher's on the old map bit for bit, and that result meets the oracle on
    the old map.  After the swap, alignments from starts the oracle converges
    from on the new map (the true pose plus N(0, 5 cm) on x, y, z) move from
    their start, meet the oracle on the new map and equal a fresh matcher's."""
    from cuda_ndt_matcher_amd import NdtScanMatcher, NdtScanMatcherConfig, synth
    from cuda_ndt_matcher_amd.dual import DualMatcher

    small = synth.make_workload("small", 0)
    big = full_wl
    cfg = NdtScanMatcherConfig(resolution=2.0, max_iterations=30, trans_epsilon=0.01)
    p = oracle_mod.params()
    _case("dual matcher: background 5M-point build during alignments")
    ref_old = NdtScanMatcher(cfg)
    ref_old.set_target(small.map_points)
    # align_batch does not swap (dual_ndt_manager.rs:314-321), so every call in
    # the loop runs on the old map
    want_old = ref_old.align_batch(small.scan, [small.guess])[0]
    ref_old.close()
    o_old = oracle_mod.Grid(small.map_points, 2.0).align(small.scan, small.guess, p)
    why = parity.mismatch(want_old, o_old)
    assert why is None, f"old map: {why}"
    d = DualMatcher(cfg)
    d.set_target(small.map_points)
    assert d.start_background_update(big.map_points)
    during = 0
    t0 = time.perf_counter()
    while d.is_update_in_progress() and time.perf_counter() - t0 < 60:
        r = d.align_batch(small.scan, [small.guess])[0]
        if d.is_update_in_progress():
            during += 1
        assert np.array_equal(r.pose_vector, want_old.pose_vector) and r.iterations == want_old.iterations
    print(f"[case] {during} alignments finished while the rebuild ran", file=sys.stderr, flush=True)
    d.wait_for_update(60)
    st = d.get_status()
    assert st.swap_count == 1 and not st.in_progress
    ref_new = NdtScanMatcher(cfg)
    ref_new.set_target(big.map_points)
    for k in range(3):
        rng = np.random.default_rng(400 + k)
        start = np.asarray(big.true_pose, np.float64) + np.concatenate([rng.normal(0, 0.05, 3), np.zeros(3)])
        _case(f"dual matcher: after the swap, converging start {k}")
        got = d.align(big.scan, start)
        o = full_oracle.align(big.scan, start, p)
        assert o.status == 0, f"start {k} does not converge in the oracle"
        why = parity.mismatch(got, o)
        assert why is None, f"new map, start {k}: {why}"
        assert np.abs(np.asarray(got.pose_vector[:3]) - start[:3]).max() > 1e-3, "the alignment did not move"
        want_new = ref_new.align(big.scan, start)
        assert np.array_equal(got.pose_vector, want_new.pose_vector) and got.iterations == want_new.iterations
    ref_new.close()
    d.close()
