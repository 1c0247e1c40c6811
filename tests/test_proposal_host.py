"""Host logic of the learned initial-pose proposal (CPU only).

The device parts (descriptor, network, alignment) are replaced by stubs
here; the GPU tests cover them against their references.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest


def test_frame_round_trip_and_units():
    from cuda_ndt_matcher_amd.proposal import Frame

    f = Frame(10.0, -20.0, 4.0)
    x, y, yaw = np.array([10.0, 14.0, -6.0]), np.array([-20.0, -12.0, 0.0]), np.array([0.5, -3.0, 3.1])
    enc = f.encode(x, y, yaw)
    np.testing.assert_allclose(enc[:, 0], [0.0, 1.0, -4.0])
    np.testing.assert_allclose(enc[:, 1], [0.0, 2.0, 5.0])
    out = np.concatenate([enc, np.full((3, 1), 0.25, np.float32)], 1).reshape(1, -1)
    dec = f.decode(out)[0]
    np.testing.assert_allclose(dec[:, 0], x, atol=1e-5)
    np.testing.assert_allclose(dec[:, 1], y, atol=1e-5)
    np.testing.assert_allclose(dec[:, 2], yaw, atol=1e-6)
    assert np.all(dec[:, 3] == 0.25)


def test_wrap_range():
    from cuda_ndt_matcher_amd.proposal import wrap

    a = np.linspace(-20, 20, 4001)
    w = wrap(a)
    assert np.all(w >= -math.pi) and np.all(w < math.pi)
    np.testing.assert_allclose(np.cos(w), np.cos(a), atol=1e-12)
    np.testing.assert_allclose(np.sin(w), np.sin(a), atol=1e-12)


def test_rotated_scan_relabel_is_the_same_place():
    """A scan turned by d about the sensor's z axis and labelled yaw - d maps
    to the same map points (the training relabel)."""
    rng = np.random.default_rng(3)
    p = rng.uniform(-30, 30, (200, 3))
    t, yaw, d = np.array([5.0, -7.0, 1.8]), 0.7, -2.1

    def rz(a):
        return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])

    world = p @ rz(yaw).T + t
    turned = p @ rz(d).T
    np.testing.assert_allclose(turned @ rz(yaw - d).T + t, world, atol=1e-9)


def test_store_crop_keeps_every_binned_point():
    """Descriptors of the cropped scan equal those of the whole scan at any
    heading (the crop drops only points no rotation can bin)."""
    from cuda_ndt_matcher_amd.proposal import ScanStore
    from oracle import mlp_ref as R

    rng = np.random.default_rng(5)
    scan = np.concatenate([rng.uniform([-60, -60, -4], [60, 60, 20], (6000, 3)),
                           rng.uniform([-46, -46, -2.1], [46, 46, 14.1], (3000, 3))]).astype(np.float32)
    st = ScanStore()
    kept = st.crop(scan)
    assert 0 < kept.shape[0] < scan.shape[0]
    for t in np.linspace(-math.pi, math.pi, 9):
        rot = (np.float32(math.cos(t)), np.float32(math.sin(t)))
        assert np.array_equal(R.bev(scan, rot=rot), R.bev(kept, rot=rot))


class _Res(SimpleNamespace):
    pass


class _Matcher:
    """align_batch stub: results chosen by the test; records the poses."""

    def __init__(self, rounds):
        self.rounds = list(rounds)
        self.calls = []

    def align_batch(self, scan, poses):
        self.calls.append([np.asarray(p, float) for p in poses])
        make = self.rounds.pop(0)
        return [make(k, np.asarray(p, float)) for k, p in enumerate(poses)]


def _sample():
    from cuda_ndt_matcher_amd.proposal import Sample

    return Sample(np.zeros((10, 3), np.float32), np.zeros((10, 3), np.float32),
                  np.array([1.0, 2.0, 1.8, 0.0, 0.0, 0.3]))


def _fixed_hypotheses(monkeypatch, k=16):
    from cuda_ndt_matcher_amd import proposal

    hyp = np.stack([np.arange(k, dtype=float), -np.arange(k, dtype=float), np.linspace(-1, 1, k)], -1)[None]
    monkeypatch.setattr(proposal, "propose", lambda net, frame, bev, kk: hyp[:, :kk])
    return hyp[0]


def test_learned_search_takes_the_best_converged_hypothesis(monkeypatch):
    from cuda_ndt_matcher_amd import proposal

    hyp = _fixed_hypotheses(monkeypatch)
    m = _Matcher([lambda k, p: _Res(nvtl=2.4 + 0.01 * k, converged=True, pose_vector=p)])
    fb = []
    r, used = proposal.learned_search(m, None, None, _sample(), None, fallback=lambda s: fb.append(s))
    assert not used and not fb and len(m.calls) == 1
    # height from the prior, roll/pitch zero, (x, y, yaw) from the hypotheses
    sent = np.array(m.calls[0])
    np.testing.assert_allclose(sent[:, [0, 1, 5]], hyp)
    assert np.all(sent[:, 2] == 1.8) and np.all(sent[:, 3:5] == 0.0)
    assert r.nvtl == pytest.approx(2.55)


def test_learned_search_polishes_then_falls_back(monkeypatch):
    from cuda_ndt_matcher_amd import proposal

    _fixed_hypotheses(monkeypatch)
    first = lambda k, p: _Res(nvtl=1.0 + 0.01 * k, converged=False, pose_vector=p + 0.5)  # noqa: E731
    again = lambda k, p: _Res(nvtl=1.5, converged=True, pose_vector=p)  # noqa: E731
    m = _Matcher([first, again])
    sentinel = _Res(nvtl=3.0)
    r, used = proposal.learned_search(m, None, None, _sample(), None, polish=4, fallback=lambda s: sentinel)
    assert used and r is sentinel
    # the polish pass restarts the four best (highest NVTL) from where they stopped
    assert len(m.calls) == 2 and len(m.calls[1]) == 4
    best4 = [np.concatenate([[i, -i, 1.8, 0, 0, -1 + 2 * i / 15]]) + 0.5 for i in (15, 14, 13, 12)]
    np.testing.assert_allclose(np.array(m.calls[1]), np.array(best4))


def test_learned_search_keeps_a_polished_pose_above_threshold(monkeypatch):
    from cuda_ndt_matcher_amd import proposal

    _fixed_hypotheses(monkeypatch)
    m = _Matcher([lambda k, p: _Res(nvtl=1.0, converged=False, pose_vector=p),
                  lambda k, p: _Res(nvtl=2.6 - 0.1 * k, converged=True, pose_vector=p)])
    r, used = proposal.learned_search(m, None, None, _sample(), None, fallback=lambda s: pytest.fail("fallback"))
    assert not used and r.nvtl == pytest.approx(2.6)


def test_plain_search_particles():
    from cuda_ndt_matcher_amd import proposal

    m = _Matcher([lambda k, p: _Res(nvtl=float(k == 17), converged=True, pose_vector=p)])
    s = _sample()
    r = proposal.plain_search(m, s, 200, seed=9)
    sent = np.array(m.calls[0])
    assert sent.shape == (200, 6)
    d = sent - s.true_pose
    assert np.all(np.abs(d[:, :2]) <= 2.0) and np.all(d[:, 2] == 0.0)
    assert np.all(np.abs(d[:, 5]) <= math.pi) and np.abs(d[:, 3:5]).max() < math.radians(6)
    assert r.nvtl == 1.0 and np.array_equal(r.pose_vector, sent[17])
    # seeded
    m2 = _Matcher([lambda k, p: _Res(nvtl=0.0, converged=True, pose_vector=p)])
    proposal.plain_search(m2, s, 200, seed=9)
    np.testing.assert_array_equal(np.array(m2.calls[0]), sent)


def test_plain_search_polish_pass():
    """With polish, the four best particles (by NVTL) restart from where they
    stopped when the best has not converged; the best of both launches wins.
    A converged best needs no second launch."""
    from cuda_ndt_matcher_amd import proposal

    first = lambda k, p: _Res(nvtl=0.01 * k, converged=False, pose_vector=p + 1.0)  # noqa: E731
    again = lambda k, p: _Res(nvtl=5.0 - k, converged=True, pose_vector=p)  # noqa: E731
    m = _Matcher([first, again])
    r = proposal.plain_search(m, _sample(), 20, seed=3, polish=4)
    sent = np.array(m.calls[0])
    assert len(m.calls) == 2
    np.testing.assert_allclose(np.array(m.calls[1]), sent[[19, 18, 17, 16]] + 1.0)
    assert r.nvtl == 5.0 and np.array_equal(r.pose_vector, sent[19] + 1.0)
    m2 = _Matcher([lambda k, p: _Res(nvtl=0.01 * k, converged=True, pose_vector=p)])
    proposal.plain_search(m2, _sample(), 20, seed=3, polish=4)
    assert len(m2.calls) == 1
