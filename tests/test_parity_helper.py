"""The parity checker (oracle/parity.py) on the CPU: results taken from the
oracle itself pass, and planted mistakes (a moved pose on a stable start, a
wrong score or status on a chaotic one) are caught."""
import copy
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import parity


def _as_gpu(o, **change):
    r = SimpleNamespace(status=parity.STATUS[o.status], iterations=o.iterations, pose_vector=np.array(o.pose),
                        nvtl=o.nvtl, transform_probability=o.transform_probability,
                        num_correspondences=o.num_correspondences, oscillation_count=o.oscillation_count,
                        hessian=np.array(o.hessian), score=o.score, converged=o.status == 0)
    for k, v in change.items():
        setattr(r, k, v)
    return r


@pytest.fixture(scope="module")
def small(oracle_mod):
    from cuda_ndt_matcher_amd import synth

    w = synth.make_workload("small", 0)
    return w, oracle_mod.Grid(w.map_points, 2.0)


def _batch_starts(w):
    rng = np.random.default_rng(21)
    return [w.true_pose + np.array([rng.normal(0, 0.3), rng.normal(0, 0.3), 0, 0, 0, rng.normal(0, 0.03)])
            for _ in range(16)]


def test_oracle_result_matches_itself(small, oracle_mod):
    w, og = small
    p = oracle_mod.params()
    o = og.align(w.scan, w.guess, p)
    assert parity.check_alignment(_as_gpu(o), og, w.scan, w.guess, p, o=o)[0] == "match"


def test_moved_pose_on_a_stable_start_fails(small, oracle_mod):
    w, og = small
    p = oracle_mod.params()
    o = og.align(w.scan, w.guess, p)
    assert o.status == 0
    moved = np.array(o.pose) + np.array([0.01, 0, 0, 0, 0, 0])
    with pytest.raises(AssertionError):
        parity.check_alignment(_as_gpu(o, pose_vector=moved), og, w.scan, w.guess, p, o=o)


def test_chaotic_start_checks_score_and_status(small, oracle_mod):
    """A start of the batch test whose oracle trajectory separates under the
    probe: another perturbed run of the oracle stands in for the GPU's
    trajectory and is accepted; the same result with a score that is not
    the oracle's at its pose, or with a status no probe run shows, is not."""
    w, og = small
    p = oracle_mod.params()
    for start in _batch_starts(w):
        pr = parity.Probe(og, w.scan, start, p, 1e-10)
        if pr.chaotic and pr.statuses == {1}:
            break
    else:
        pytest.fail("no chaotic start among the batch starts")
    q = copy.copy(p)
    q.noise, q.noise_seed = 1e-10, 17
    other = og.align(w.scan, start, q)
    kind, _ = parity.check_alignment(_as_gpu(other), og, w.scan, start, p, o=pr.o)
    assert kind in ("match", "chaotic")
    with pytest.raises(AssertionError):
        parity.check_alignment(_as_gpu(other, score=other.score * (1 + 1e-6)), og, w.scan, start, p, o=pr.o)
    with pytest.raises(AssertionError):
        parity.check_alignment(_as_gpu(other, status="SingularHessian"), og, w.scan, start, p, o=pr.o)
