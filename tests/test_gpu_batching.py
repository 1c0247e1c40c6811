"""Graph replay and the autotuner (GPU): both must leave the results alone.

A graph replay runs the same kernel on the same inputs as a direct launch,
so with the same workgroup count the results are bit-identical.  The
autotuner only chooses the workgroup count; its answer equals a direct run
at the count it chose.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def small():
    from cuda_ndt_matcher_amd import synth

    return synth.make_workload("small", 0)


def _matcher(points, **cfg):
    from cuda_ndt_matcher_amd import NdtScanMatcher, NdtScanMatcherConfig

    m = NdtScanMatcher(NdtScanMatcherConfig(**cfg))
    m.set_target(points)
    return m


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.pose_vector, y.pose_vector)
        assert x.iterations == y.iterations and x.status == y.status
        for a, b in ((x.score, y.score), (x.nvtl, y.nvtl), (x.transform_probability, y.transform_probability)):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(x.hessian, y.hessian, equal_nan=True)
        assert x.num_correspondences == y.num_correspondences


def _poses(small, seed, m):
    rng = np.random.default_rng(seed)
    return [small.true_pose + np.array([rng.normal(0, 0.5), rng.normal(0, 0.5), 0, 0, 0, rng.normal(0, 0.05)])
            for _ in range(m)]


@pytest.mark.parametrize("wg", [1, 4])
def test_graph_replays_match_direct_launches(small, wg):
    """Several replays with new poses (host work and other launches between
    them), then a new batch size (the graph is rebuilt)."""
    import torch

    _print_case(f"graph batch wg={wg}")
    direct = _matcher(small.map_points, workgroups_per_alignment=wg)
    graph = _matcher(small.map_points, workgroups_per_alignment=wg, use_graph=True)
    for rep, m in enumerate((16, 16, 16, 8, 16)):
        poses = _poses(small, 100 + rep, m)
        want = direct.align_batch(small.scan, poses)
        got = graph.align_batch(small.scan, poses)
        assert graph.timing()["align_graph"] == 1
        _same(got, want)
        # unrelated device work between replays
        graph.evaluate_nvtl(small.scan, poses[0])
        torch.ones(1024, device="cuda:0").sum().item()


def test_graph_single_alignment_cooperative(small):
    """One alignment over many cooperating workgroups (cooperative graph node)."""
    _print_case("graph single")
    direct = _matcher(small.map_points)
    graph = _matcher(small.map_points, use_graph=True)
    for k in range(3):
        pose = _poses(small, 200 + k, 1)[0]
        want = direct.align(small.scan, pose)
        got = graph.align(small.scan, pose)
        t = graph.timing()
        assert t["align_graph"] == 1 and t["align_workgroups"] > 1
        _same([got], [want])


def test_graph_follows_a_new_map(small):
    """set_target after a replay: the next call rebuilds against the new grid."""
    from cuda_ndt_matcher_amd import synth

    _print_case("graph new map")
    other = synth.make_workload("small", 3)
    graph = _matcher(small.map_points, use_graph=True, workgroups_per_alignment=2)
    graph.align_batch(small.scan, _poses(small, 300, 4))
    graph.set_target(other.map_points)
    direct = _matcher(other.map_points, workgroups_per_alignment=2)
    poses = [other.true_pose + d for d in np.random.default_rng(5).normal(0, 0.2, (4, 6)) * [1, 1, 0, 0, 0, 0.1]]
    _same(graph.align_batch(other.scan, poses), direct.align_batch(other.scan, poses))


def test_autotune_picks_a_count_and_keeps_it(small):
    _print_case("autotune")
    tuned = _matcher(small.map_points, autotune=True)
    poses = _poses(small, 400, 8)
    got = tuned.align_batch(small.scan, poses)
    table = tuned.tuning()
    assert len(table) == 1
    e = table[0]
    assert e["scan_bits"] == int(small.scan.shape[0]).bit_length() and e["batch_bits"] == (8).bit_length()
    W = e["workgroups_per_alignment"]
    assert W >= 1 and e["kernel_ms"] > 0
    t = tuned.timing()
    assert t["align_tuned"] == 1 and t["align_workgroups"] == 8 * W
    _same(got, _matcher(small.map_points, workgroups_per_alignment=W).align_batch(small.scan, poses))
    # same bucket: no new entry, same count
    tuned.align_batch(small.scan, _poses(small, 401, 8))
    assert tuned.tuning() == table and tuned.timing()["align_workgroups"] == 8 * W
    # a new bucket gets its own entry
    tuned.align(small.scan, poses[0])
    assert len(tuned.tuning()) == 2


def test_phase_profile_on_the_first_call(small):
    """The phase buffer exists before the first launch reads its address."""
    _print_case("phase profile first call")
    m = _matcher(small.map_points, profile_phases=True)
    r = m.align(small.scan, small.guess)
    ph = m.phase_profile()
    assert ph["iterations"] == r.iterations and ph["total_us"] > 0 and ph["derivative_us"] > 0


def test_batch_beyond_coresident_capacity(small):
    """More alignments than workgroups that fit on the GPU at once (one
    workgroup each, no cross-workgroup barrier): every result equals the
    same alignment run alone with one workgroup."""
    _print_case("batch of 600")
    m = _matcher(small.map_points, workgroups_per_alignment=1)
    poses = _poses(small, 500, 600)
    res = m.align_batch(small.scan, poses)
    assert m.timing()["align_workgroups"] == 600
    for k in (0, 257, 599):
        _same([res[k]], [m.align(small.scan, poses[k])])


def test_tiny_scans(small):
    """One- and few-point scans align (or report no correspondences) without
    error, and the multi-scan path agrees with single alignments."""
    _print_case("tiny scans")
    m = _matcher(small.map_points, workgroups_per_alignment=2)
    scans = [(small.scan[:1], small.guess), (small.scan[:3], small.guess), (small.scan[:17], small.guess)]
    many = m.align_parallel_scans(scans)
    for (s, p), r in zip(scans, many):
        alone = m.align(s, p)
        _same([r], [alone])
        assert r.status in ("Converged", "MaxIterations", "NoCorrespondences", "SingularHessian")
