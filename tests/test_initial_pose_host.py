"""Initial-pose estimation host logic (CPU): the tree-structured Parzen
estimator and the startup/guided particle flow, with the alignments stubbed.
Known answers from the reference's own tests (tpe.rs:264-302,
initial_pose.rs:496-538) plus closed forms."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from cuda_ndt_matcher_amd import initial_pose as IP


def test_log_sum_exp_known_answers():
    # tpe.rs:287-296: log(e^1 + e^2 + e^3) ~ 3.407; empty -> -inf
    v = IP.log_sum_exp([1.0, 2.0, 3.0])
    assert abs(v - 3.407) < 0.01
    assert v == pytest.approx(math.log(math.e + math.e ** 2 + math.e ** 3), rel=1e-15)
    assert IP.log_sum_exp([]) == -math.inf
    assert IP.log_sum_exp([-math.inf, -math.inf]) == -math.inf
    # large values do not overflow
    assert IP.log_sum_exp([1000.0, 1000.0]) == pytest.approx(1000.0 + math.log(2.0))


def test_above_num_and_startup_phase():
    t = IP.TreeStructuredParzenEstimator(10, np.zeros(6), [1.0, 1.0, 0.5, 0.1, 0.1, math.pi],
                                         np.random.default_rng(0))
    assert t.compute_above_num() == 0
    for i in range(10):  # tpe.rs:264-279: the first ten are random draws
        x = t.get_next_input()
        assert -math.pi <= x[5] < math.pi
        t.add_trial(x, float(i))
    assert [tr.score for tr in t.trials] == sorted([float(i) for i in range(10)], reverse=True)
    for n, want in ((1, 1), (9, 1), (25, 2), (99, 9), (100, 10), (500, 10)):
        t.trials = [IP.Trial(np.zeros(6), 0.0)] * n
        assert t.compute_above_num() == want


def test_random_input_follows_the_prior():
    mean = np.array([10.0, -5.0, 2.0, 0.05, -0.02, 1.0])
    sd = np.array([1.0, 2.0, 0.1, 0.01, 0.02, 0.3])
    t = IP.TreeStructuredParzenEstimator(100, mean, sd, np.random.default_rng(1))
    xs = np.array([t.generate_random_input() for _ in range(20000)])
    assert np.all(np.abs(xs[:, :5].mean(0) - mean[:5]) <= 4 * sd[:5] / math.sqrt(20000))
    np.testing.assert_allclose(xs[:, :5].std(0), sd[:5], rtol=0.03)
    # yaw ignores the prior: uniform over [-pi, pi)
    assert xs[:, 5].min() >= -math.pi and xs[:, 5].max() < math.pi
    assert abs(xs[:, 5].mean()) < 0.05 and xs[:, 5].std() == pytest.approx(math.pi / math.sqrt(3), rel=0.02)
    # an invalid deviation falls back to N(0, 1)
    t2 = IP.TreeStructuredParzenEstimator(1, [5.0] * 6, [-1.0] * 6, np.random.default_rng(2))
    ys = np.array([t2.generate_random_input() for _ in range(5000)])
    assert abs(ys[:, 0].mean()) < 0.1 and ys[:, 0].std() == pytest.approx(1.0, rel=0.05)


def test_log_likelihood_ratio_closed_form():
    t = IP.TreeStructuredParzenEstimator(0, np.zeros(6), np.ones(6), np.random.default_rng(3))
    a = np.array([1.0, 2.0, 9.0, 0.0, 0.0, 3.1])
    b = np.array([-1.0, 0.5, -9.0, 0.0, 0.0, -3.0])
    c = np.array([0.3, 0.0, 0.0, 0.0, 0.0, 0.2])
    t.add_trial(a, 2.0)
    t.add_trial(b, 1.0)
    t.add_trial(c, 0.5)
    x = np.array([0.9, 1.7, 100.0, 5.0, 5.0, -3.1])
    s = IP.BASE_STDDEV

    def lp(u, v):
        dy = (u[5] - v[5] + math.pi) % (2 * math.pi) - math.pi  # the wrap, written differently
        return sum(-0.5 * (d / s[i]) ** 2 - math.log(s[i]) for d, i in ((u[0] - v[0], 0), (u[1] - v[1], 1), (dy, 5)))

    # z, roll and pitch play no part; yaw distance wraps (3.1 vs -3.1 is 0.083 rad)
    assert t.log_gaussian_pdf(x, a) == pytest.approx(lp(x, a), rel=1e-12)
    above = lp(x, a) + math.log(1.0)
    below = math.log(math.exp(lp(x, b) + math.log(0.5)) + math.exp(lp(x, c) + math.log(0.5)))
    assert t.log_likelihood_ratio(x, 1) == pytest.approx(above - 5.0 * below, rel=1e-12)
    assert t.log_likelihood_ratio(x, 3) == 0.0  # above_num >= n


def test_guided_input_prefers_good_trials():
    """After startup the proposal lands near the best trial rather than the
    others (x, y, yaw kernels)."""
    t = IP.TreeStructuredParzenEstimator(5, [0, 0, 0, 0, 0, 0], [3.0, 3.0, 0.1, 0.01, 0.01, 0.0],
                                         np.random.default_rng(4))
    good = np.array([2.0, -2.0, 0, 0, 0, 1.0])
    t.add_trial(good, 10.0)
    for k in range(19):
        t.add_trial(np.array([-2.0 + 0.1 * k, 2.0, 0, 0, 0, -2.0]), 0.0)
    near = 0
    for _ in range(30):
        x = t.get_next_input()
        near += np.hypot(*(x[:2] - good[:2])) < np.hypot(*(x[:2] - np.array([-1.0, 2.0])))
    assert near >= 25


def test_batched_proposals_match_single_proposals():
    """get_next_inputs(k) makes k proposals from the same trials: each is the
    first candidate with the largest ratio among its own 100 candidates (its
    fallback draw first), so with the same generator state one batched
    proposal equals one single proposal; trials whose ratios are all NaN
    give the fallback draw."""
    def fresh(seed):
        t = IP.TreeStructuredParzenEstimator(5, np.zeros(6), [2.0, 2.0, 0.1, 0.01, 0.01, 0.0],
                                             np.random.default_rng(seed))
        for k in range(30):
            t.add_trial(np.array([0.2 * k - 3.0, 1.0 - 0.1 * k, 0, 0, 0, 0.1 * k]), float(k % 7))
        return t

    a, b = fresh(11), fresh(11)
    np.testing.assert_array_equal(a.get_next_inputs(1)[0], b.get_next_input())
    # a batch of 6 against the brute-force argmax over the same draws
    t, ref = fresh(12), fresh(12)
    got = t.get_next_inputs(6)
    draws = ref.random_inputs(6 * 101).reshape(6, 101, 6)
    for i in range(6):
        ratios = [ref.log_likelihood_ratio(c, ref.compute_above_num()) for c in draws[i, 1:]]
        np.testing.assert_array_equal(got[i], draws[i, 1 + int(np.argmax(ratios))])
    # NaN trial inputs: every ratio is NaN, the fallback draw is returned
    n, r = fresh(13), fresh(13)
    for tr in n.trials:
        tr.input = np.full(6, np.nan)
    n._rows = None
    got = n.get_next_inputs(3)
    draws = r.random_inputs(3 * 101).reshape(3, 101, 6)
    np.testing.assert_array_equal(got, draws[:, 0])
    # before the startup trials are in, proposals are prior draws
    s0 = IP.TreeStructuredParzenEstimator(100, np.zeros(6), np.ones(6), np.random.default_rng(14))
    s1 = IP.TreeStructuredParzenEstimator(100, np.zeros(6), np.ones(6), np.random.default_rng(14))
    np.testing.assert_array_equal(s0.get_next_inputs(4), s1.random_inputs(4))


def test_best_particle_takes_the_last_of_equals():
    P = IP.Particle
    ps = [P(np.zeros(6), np.full(6, k), s, 1) for k, s in enumerate([1.0, 3.0, 2.0, 3.0, float("nan")])]
    # NaN compares equal, so the last element (k = 4) wins after the tie at 3.0
    assert IP.select_best_particle(ps).result_pose[0] == 4
    assert IP.select_best_particle(ps[:4]).result_pose[0] == 3
    assert IP.select_best_particle([]) is None


def test_prior_stddev_floors():
    cov = np.zeros(36)
    cov[0], cov[7], cov[14], cov[21], cov[28], cov[35] = 4.0, 0.0001, 0.25, 1e-6, 0.04, 9.0
    np.testing.assert_allclose(IP.prior_stddev(cov), [2.0, 0.1, 0.5, 0.01, 0.2, 3.0])


class _Stub:
    """align/align_batch stub: the result sits 0.1 from the start; score and
    NVTL depend on the start's x."""

    def __init__(self):
        self.batches = []
        self.singles = 0

    def _one(self, p):
        p = np.asarray(p, float)
        x = p[0]
        return SimpleNamespace(pose_vector=p + 0.1, score=20.0 + x, nvtl=max(0.0, 2.0 + x), iterations=7)

    def align_batch(self, scan, poses):
        self.batches.append(len(poses))
        return [self._one(p) for p in poses]

    def align(self, scan, pose):
        self.singles += 1
        return self._one(pose)


@pytest.mark.parametrize("guided_batch", [1, 10])
def test_estimate_initial_pose_flow(guided_batch):
    m = _Stub()
    cov = np.diag([1.0, 1.0, 0.01, 1e-4, 1e-4, 1.0]).reshape(-1)
    r = IP.estimate_initial_pose(m, None, [0, 0, 1.8, 0, 0, 0], cov, particles_num=200, n_startup_trials=100,
                                 guided_batch=guided_batch, seed=5)
    assert len(r.particles) == 200
    if guided_batch == 1:
        assert m.batches == [100] and m.singles == 100
    else:
        assert m.batches == [100] + [10] * 10 and m.singles == 0
    for p in r.particles:
        x = p.initial_pose[0]
        want = 2.0 + x if 2.0 + x > 0 else math.exp(-(20.0 + x) / 10.0) * 5.0
        assert p.score == pytest.approx(want)
        np.testing.assert_allclose(p.result_pose, p.initial_pose + 0.1)
    best = max(p.score for p in r.particles)
    assert r.score == best and r.reliable == (best >= 2.3)
    np.testing.assert_allclose(r.pose, [p for p in r.particles if p.score == best][-1].result_pose)
    # the estimator saw every particle, scored exp(-score / 10) and read back as Euler angles
    assert len(r.particles) == 200


def test_estimate_initial_pose_is_seeded():
    a = IP.estimate_initial_pose(_Stub(), None, [0] * 6, np.eye(6).reshape(-1), 30, 10, seed=9)
    b = IP.estimate_initial_pose(_Stub(), None, [0] * 6, np.eye(6).reshape(-1), 30, 10, seed=9)
    np.testing.assert_array_equal(np.array([p.initial_pose for p in a.particles]),
                                  np.array([p.initial_pose for p in b.particles]))
