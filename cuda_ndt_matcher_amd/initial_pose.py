"""Initial-pose estimation with a tree-structured Parzen estimator.

Follows cuda_ndt_matcher/src/initial_pose.rs:115-400 (estimate_initial_pose),
tpe.rs:41-245 (TreeStructuredParzenEstimator) and particle.rs:36-42
(select_best_particle):

* particles are drawn around a prior pose: x, y, z, roll, pitch from normals
  whose deviations come from the prior's covariance (floored at 0.1 m and
  0.01 rad), yaw uniform in [-pi, pi);
* the first n_startup_trials particles are aligned in one batched launch;
* the rest are proposed by the estimator one at a time: of 100 random
  candidates, the one with the best log-likelihood ratio between the best
  tenth of the trials so far (at most 10) and the others (weighted 5x) on
  x, y and yaw, with fixed kernel widths 0.25 m and 2.5 degrees;
* each aligned particle is scored for the estimator with exp(-score / 10) of
  the alignment's NDT score, and for the final choice with its NVTL (or that
  value times 5 when the NVTL is 0); the highest choice score wins (the last
  one among equals), and the result is reliable at NVTL >= 2.3.

The estimator's score keeps the reference's form: exp(-score / 10) of an NDT
score that is higher for better alignments, so the guided phase leans
towards the trials that aligned worse; the final choice by NVTL is not
affected.  Randomness comes from a seeded generator (the reference uses a
thread-local one).  With guided_batch = k > 1 the guided phase proposes k
particles per round from the same trials and aligns them in one launch
(guided_batch = 1 is the reference's order).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from functools import cmp_to_key
from typing import List, Optional, Sequence

import numpy as np

TRANS_X, TRANS_Y, TRANS_Z, ANGLE_X, ANGLE_Y, ANGLE_Z = range(6)
INPUT_DIM = 6
# kernel widths of the estimator (tpe.rs:62-70)
BASE_STDDEV = (0.25, 0.25, 0.25, math.radians(1.0), math.radians(1.0), math.radians(2.5))
N_EI_CANDIDATES = 100
NVTL_THRESHOLD = 2.3


def log_sum_exp(values: Sequence[float]) -> float:
    """log(sum(exp(v))) with the largest term factored out; -inf when empty."""
    if len(values) == 0:
        return -math.inf
    m = max(values)
    if math.isinf(m):
        return m
    return m + math.log(sum(math.exp(v - m) for v in values))


def log_gaussian_pdf_1d(diff: float, sigma: float) -> float:
    """Unnormalised log density: -(diff / sigma)^2 / 2 - ln sigma."""
    return -0.5 * (diff / sigma) ** 2 - math.log(sigma)


def _by_score_desc(a, b) -> int:
    # partial_cmp(b, a) with unordered (NaN) pairs equal; the sort is stable
    if a.score != a.score or b.score != b.score:
        return 0
    return -1 if a.score > b.score else (1 if a.score < b.score else 0)


@dataclass
class Trial:
    input: np.ndarray
    score: float


class TreeStructuredParzenEstimator:
    """The estimator of tpe.rs (direction: maximise)."""

    def __init__(self, n_startup_trials: int, mean: Sequence[float], stddev: Sequence[float],
                 rng: Optional[np.random.Generator] = None):
        self.n_startup_trials = int(n_startup_trials)
        self.mean = np.asarray(mean, np.float64)
        self.stddev = np.asarray(stddev, np.float64)
        self.base_stddev = np.asarray(BASE_STDDEV)
        self.trials: List[Trial] = []
        self.rng = rng if rng is not None else np.random.default_rng()
        self._rows = None  # trial inputs (x, y, yaw) in rank order

    def add_trial(self, input, score: float) -> None:
        self.trials.append(Trial(np.asarray(input, np.float64), float(score)))
        self.trials.sort(key=cmp_to_key(_by_score_desc))
        self._rows = None

    def compute_above_num(self) -> int:
        n = len(self.trials)
        return 0 if n == 0 else min(max(n // 10, 1), 10)

    def get_next_input(self) -> np.ndarray:
        above = self.compute_above_num()
        if len(self.trials) < self.n_startup_trials or above == 0:
            return self.generate_random_input()
        return self.generate_tpe_input(above)

    def random_inputs(self, count: int) -> np.ndarray:
        """`count` draws from the prior: x, y, z, roll, pitch normal (N(0, 1)
        where the deviation or mean is not a valid number, as Normal::new
        failing does), yaw uniform in [-pi, pi)."""
        m, sd = self.mean[:ANGLE_Z].copy(), self.stddev[:ANGLE_Z].copy()
        bad = ~(np.isfinite(sd) & np.isfinite(m) & (sd >= 0.0))
        m[bad], sd[bad] = 0.0, 1.0
        x = np.empty((count, INPUT_DIM))
        x[:, :ANGLE_Z] = self.rng.normal(m, sd, (count, ANGLE_Z))
        x[:, ANGLE_Z] = self.rng.uniform(-math.pi, math.pi, count)
        return x

    def generate_random_input(self) -> np.ndarray:
        return self.random_inputs(1)[0]

    def generate_tpe_input(self, above_num: int) -> np.ndarray:
        return self.generate_tpe_inputs(above_num, 1)[0]

    def generate_tpe_inputs(self, above_num: int, count: int) -> np.ndarray:
        """`count` independent proposals from the current trials: each draws a
        fallback and 100 candidates and keeps the first candidate with the
        largest log-likelihood ratio, or the fallback when every ratio is -inf
        or not a number."""
        per = N_EI_CANDIDATES + 1
        draws = self.random_inputs(count * per).reshape(count, per, INPUT_DIM)
        cands = draws[:, 1:].reshape(-1, INPUT_DIM)
        r = self.log_likelihood_ratios(cands, above_num).reshape(count, N_EI_CANDIDATES)
        r = np.where(np.isnan(r), -np.inf, r)
        k = np.argmax(r, axis=1)
        rows = np.arange(count)
        ok = r[rows, k] > -math.inf
        return np.where(ok[:, None], draws[rows, 1 + k], draws[:, 0])

    def get_next_inputs(self, count: int) -> np.ndarray:
        """`count` proposals from the same trials (get_next_input `count`
        times without trials added in between)."""
        above = self.compute_above_num()
        if len(self.trials) < self.n_startup_trials or above == 0:
            return self.random_inputs(count)
        return self.generate_tpe_inputs(above, count)

    def log_gaussian_pdf(self, input, mu) -> float:
        r = 0.0
        for i in (TRANS_X, TRANS_Y, ANGLE_Z):
            d = float(input[i] - mu[i])
            if i == ANGLE_Z:
                while d >= math.pi:
                    d -= 2.0 * math.pi
                while d < -math.pi:
                    d += 2.0 * math.pi
            r += log_gaussian_pdf_1d(d, self.base_stddev[i])
        return r

    def _log_pdf_rows(self, cands: np.ndarray) -> np.ndarray:
        """(candidates x trials) log densities on x, y and yaw (yaw distance
        wrapped into [-pi, pi))."""
        if self._rows is None:
            self._rows = np.array([t.input[[TRANS_X, TRANS_Y, ANGLE_Z]] for t in self.trials]).T.copy()
        sig = self.base_stddev[[TRANS_X, TRANS_Y, ANGLE_Z]]
        two_pi = 2.0 * math.pi
        acc = np.full((cands.shape[0], self._rows.shape[1]), -float(np.log(sig).sum()))
        for j, col in enumerate((TRANS_X, TRANS_Y, ANGLE_Z)):
            d = np.subtract.outer(cands[:, col], self._rows[j])
            if col == ANGLE_Z:
                d -= two_pi * np.floor((d + math.pi) / two_pi)
            d *= 1.0 / sig[j]
            d *= d
            acc -= 0.5 * d
        return acc

    def log_likelihood_ratios(self, cands: np.ndarray, above_num: int) -> np.ndarray:
        n = len(self.trials)
        if n == 0 or above_num == 0 or above_num >= n:
            return np.zeros(len(cands))
        lp = self._log_pdf_rows(np.atleast_2d(cands))
        above = _log_sum_exp_rows(lp[:, :above_num]) + math.log(1.0 / above_num)
        below = _log_sum_exp_rows(lp[:, above_num:]) + math.log(1.0 / (n - above_num))
        return above - below * 5.0

    def log_likelihood_ratio(self, input, above_num: int) -> float:
        return float(self.log_likelihood_ratios(np.asarray(input, np.float64)[None, :], above_num)[0])


def _log_sum_exp_rows(v: np.ndarray) -> np.ndarray:
    """Row-wise log_sum_exp.  Terms more than 700 below their row's maximum
    are taken at exp(-700): the maximum contributes 1 to the sum, so they
    cannot change it, and exp of arguments below about -708 (subnormal
    results) runs tens of times slower."""
    m = v.max(axis=1)
    fin = np.isfinite(m)
    out = m.copy()
    if fin.any():
        d = v[fin] - m[fin, None]
        np.maximum(d, -700.0, out=d)
        out[fin] = m[fin] + np.log(np.exp(d).sum(axis=1))
    return out


@dataclass
class Particle:
    initial_pose: np.ndarray  # [x, y, z, roll, pitch, yaw]
    result_pose: np.ndarray
    score: float
    iterations: int


@dataclass
class InitialPoseResult:
    pose: np.ndarray
    score: float
    reliable: bool
    particles: List[Particle] = field(default_factory=list)


def select_best_particle(particles: Sequence[Particle]) -> Optional[Particle]:
    """Highest score; the last one among equals (Iterator::max_by)."""
    best = None
    for p in particles:
        if best is None or _by_score_desc(best, p) >= 0:  # best.score <= p.score or unordered
            best = p
    return best


def prior_stddev(covariance: Sequence[float]) -> np.ndarray:
    """Deviations from the diagonal of a 6x6 row-major covariance, floored
    at 0.1 (translation) and 0.01 (rotation) (initial_pose.rs:126-134)."""
    c = np.asarray(covariance, np.float64).reshape(36)
    floors = (0.1, 0.1, 0.1, 0.01, 0.01, 0.01)
    return np.array([max(math.sqrt(max(c[7 * i], 0.0)), floors[i]) for i in range(6)])


def _euler(pose_vector) -> np.ndarray:
    """The result pose as the reference reads it back: nalgebra Euler angles
    of its rotation (quaternion_to_rpy)."""
    from .isometry import Isometry3

    return Isometry3.from_pose_vector(pose_vector).to_pose_vector()


def estimate_initial_pose(matcher, scan, mean, covariance, particles_num: int = 200, n_startup_trials: int = 100,
                          guided_batch: int = 1, seed: int = 0,
                          nvtl_threshold: float = NVTL_THRESHOLD) -> InitialPoseResult:
    """estimate_initial_pose (initial_pose.rs:115-400) over `matcher`'s map.

    mean: prior pose [x, y, z, roll, pitch, yaw]; covariance: its 6x6
    covariance (36 values, row-major).  Returns the best particle's pose and
    choice score, whether it clears the NVTL threshold, and every particle."""
    tpe = TreeStructuredParzenEstimator(n_startup_trials, mean, prior_stddev(covariance),
                                        np.random.default_rng(seed))
    particles: List[Particle] = []

    def take(inputs, results):
        for x, r in zip(inputs, results):
            tp = math.exp(-r.score / 10.0)
            nvtl = r.nvtl if math.isfinite(r.nvtl) else 0.0
            particles.append(Particle(np.asarray(x), r.pose_vector, nvtl if nvtl > 0.0 else tp * 5.0,
                                      int(r.iterations)))
            tpe.add_trial(_euler(r.pose_vector), tp)

    startup = min(int(n_startup_trials), int(particles_num))
    if startup > 0:
        inputs = list(tpe.get_next_inputs(startup))
        take(inputs, matcher.align_batch(scan, inputs))
    remaining = int(particles_num) - len(particles)
    k = max(1, int(guided_batch))
    while remaining > 0:
        inputs = list(tpe.get_next_inputs(min(k, remaining)))
        results = [matcher.align(scan, inputs[0])] if len(inputs) == 1 else matcher.align_batch(scan, inputs)
        take(inputs, results)
        remaining -= len(inputs)
    best = select_best_particle(particles)
    if best is None:
        raise RuntimeError("No particles evaluated successfully")
    return InitialPoseResult(best.result_pose, best.score, best.score >= nvtl_threshold, particles)
