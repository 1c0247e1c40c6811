"""Parity checks of GPU alignments against the CPU oracle.

TEST INFRASTRUCTURE ONLY (used by tests/ and bench.py's parity rows, always
as the checker).

The bar (alignments): status and iteration count equal, pose within 1 mm /
0.01 deg, NVTL and TP within 1e-4; on converged runs also the final
correspondence count, oscillation count and Hessian.  `check_alignment` holds
a GPU result to the bar first.  Only when it misses does it run the
conditioning probe: the oracle re-run with every iteration's score, gradient
and Hessian perturbed at `noise` relative to their largest entries (five
seeds).  The default noise, 1e-10, is above the difference that another
summation order and another exponential make to an FP64 sum over tens of
thousands of points (about 1e-13 relative); FP32 per-point terms use 1e-6.

Many starts are chaotic under the reference's clamped Newton step: with an
indefinite Hessian the eigenvalues above -1e-3 are clamped to -1e-3, the step
direction follows the clamped eigenvectors, and a rounding-level difference
grows by orders of magnitude per iteration (measured on the GPU against the
oracle: last-iterate Hessians 1e-14 apart after one iteration, 1e-10 after
two, 1e-7 after three).  The probe classifies a start by its trajectory: the
last iterate's Hessian of the perturbed runs against the unperturbed one.

* Trajectory stable under the probe and the GPU misses the bar: failure.
* Trajectory chaotic: the GPU's own trajectory is one of the outcomes the
  reference's solver can produce from that start.  What can still be checked
  is that the result is what the solver reports for the trajectory it took:
  max iterations with the iteration count at its maximum; converged only at
  a pose where the
  oracle, started there, stops at its first iteration without moving; and
  the reported score and NVTL equal to the oracle's evaluation at the
  reported pose.
* Between the two (trajectory stable, best-score pose moved under the
  probe): the same checks plus the pose within twice the probe's spread.
"""
from __future__ import annotations

import copy
from concurrent.futures import ThreadPoolExecutor

import numpy as np

STATUS = ["Converged", "MaxIterations", "NoCorrespondences", "SingularHessian"]
POSE_M = 1e-3
POSE_RAD = np.deg2rad(0.01)
SCORE_TOL = 1e-4
PROBE_SEEDS = (1, 2, 3, 4, 5)
CHAOTIC_HESSIAN_REL = 1e-6   # last-iterate Hessians this far apart: the trajectories have separated


def angle_diff(a, b) -> np.ndarray:
    """Elementwise difference of angles, wrapped to [-pi, pi]."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return np.abs((d + np.pi) % (2 * np.pi) - np.pi)


def pose_errors(pv, opose):
    pv = np.asarray(pv, np.float64)
    opose = np.asarray(opose, np.float64)
    return float(np.abs(pv[:3] - opose[:3]).max()), float(angle_diff(pv[3:], opose[3:]).max())


def _hess_rel(a, b) -> float:
    b = np.asarray(b, np.float64)
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(1.0, np.abs(b).max()))


def mismatch(r, o, pose_m=POSE_M, pose_rad=POSE_RAD, score_tol=SCORE_TOL, hessian_rel=1e-6):
    """Why GPU result r misses oracle result o at the parity bar (None when it
    meets it).  At max iterations the reported Hessian and pair count belong
    to the last iterate, which the bar does not cover; on converged runs they
    are the final pose's and must agree."""
    if r.status != STATUS[o.status]:
        return f"status {r.status} vs {STATUS[o.status]}"
    if r.iterations != o.iterations:
        return f"iterations {r.iterations} vs {o.iterations}"
    dt, dr = pose_errors(r.pose_vector, o.pose)
    if dt > pose_m or dr > pose_rad:
        return f"pose {dt * 1e3:.4f} mm / {np.rad2deg(dr):.5f} deg"
    if abs(r.nvtl - o.nvtl) > score_tol:
        return f"nvtl {r.nvtl} vs {o.nvtl}"
    if abs(r.transform_probability - o.transform_probability) > score_tol:
        return f"tp {r.transform_probability} vs {o.transform_probability}"
    if o.status == 0:
        if r.num_correspondences != o.num_correspondences:
            return f"correspondences {r.num_correspondences} vs {o.num_correspondences}"
        if r.oscillation_count != o.oscillation_count:
            return f"oscillation {r.oscillation_count} vs {o.oscillation_count}"
        h = np.abs(np.asarray(r.hessian) - o.hessian).max()
        if h > hessian_rel * max(1.0, np.abs(o.hessian).max()):
            return f"hessian differs by {h:.3e}"
    return None


class Probe:
    """Oracle re-runs under the conditioning probe, relative to the
    unperturbed run o."""

    def __init__(self, og, scan, pose, params, noise, o=None, seeds=PROBE_SEEDS):
        self.o = o if o is not None else og.align(scan, pose, params)
        self.statuses, self.iterations = {self.o.status}, {self.o.iterations}
        self.spread_m = self.spread_rad = self.trajectory = 0.0
        for seed in seeds:
            p = copy.copy(params)
            p.noise = noise
            p.noise_seed = seed
            q = og.align(scan, pose, p)
            self.statuses.add(q.status)
            self.iterations.add(q.iterations)
            dt, dr = pose_errors(q.pose, self.o.pose)
            self.spread_m, self.spread_rad = max(self.spread_m, dt), max(self.spread_rad, dr)
            self.trajectory = max(self.trajectory, _hess_rel(q.hessian, self.o.hessian))

    @property
    def chaotic(self) -> bool:
        return self.trajectory > CHAOTIC_HESSIAN_REL

    @property
    def stable(self) -> bool:
        return (not self.chaotic and len(self.statuses) == 1 and len(self.iterations) == 1
                and self.spread_m <= 1e-5 and self.spread_rad <= 1e-6)


def probe(og, scan, pose, params, noise, seeds=PROBE_SEEDS):
    """(stable, spread_m, spread_rad, statuses, iterations) of the probe."""
    pr = Probe(og, scan, pose, params, noise, seeds=seeds)
    return pr.stable, pr.spread_m, pr.spread_rad, pr.statuses, pr.iterations


def oracle_score(og, scan, pose, params):
    """The solver's score at a pose: the derivative pass's score plus the
    regularisation term when it is active (solver.rs:180-190)."""
    from oracle import oracle as O

    s, _, _, corr = og.derivatives(scan, pose, params, False)
    if params.regularization_enabled and params.has_regularization_pose:
        s += O.regularization(params, pose, corr)[0]
    return s


def check_alignment(r, og, scan, pose, params, o=None, noise=1e-10, pose_m=POSE_M, pose_rad=POSE_RAD,
                    hessian_rel=1e-6, score_rel=1e-9):
    """Asserts GPU result r against the oracle (see the module docstring).
    Returns ("match" | "bounded" | "chaotic" | "stationary", probe spread in
    metres).  score_rel: how closely the GPU's reported score must equal the
    oracle's evaluation at the reported pose (FP32 terms: 2e-5)."""
    if o is None:
        o = og.align(scan, pose, params)
    why = mismatch(r, o, pose_m, pose_rad, hessian_rel=hessian_rel)
    if why is None:
        return "match", 0.0
    pr = Probe(og, scan, pose, params, noise, o=o)
    assert not pr.stable, f"oracle is stable under the probe but the GPU result differs: {why} " \
                          f"(oracle status {STATUS[o.status]}, {o.iterations} iterations, score {o.score!r}; " \
                          f"GPU score {r.score!r})"
    seen = r.status in {STATUS[s] for s in pr.statuses}
    assert r.status in ("MaxIterations", "Converged") or seen, \
        f"status {r.status} not among the probe's {pr.statuses}"
    if r.status == "MaxIterations":
        assert seen or pr.chaotic, "MaxIterations where every probe run converged"
        assert r.iterations == params.max_iterations, f"{r.iterations} iterations at max iterations"
    if r.status in ("MaxIterations", "Converged"):
        # The reported score and NVTL are the solver's at the reported pose.
        s = oracle_score(og, scan, r.pose_vector, params)
        assert abs(r.score - s) <= score_rel * max(1.0, abs(s)), f"score {r.score!r} vs oracle at that pose {s!r}"
        nv = og.nvtl(scan, r.pose_vector, params)
        assert abs(r.nvtl - nv) <= SCORE_TOL, f"nvtl {r.nvtl} vs oracle at that pose {nv}"
    if r.status == "Converged":
        # A converged result, at whatever iteration its trajectory got there,
        # must be a point the reference's solver itself accepts: started
        # there, the oracle converges at its first iteration without moving.
        assert r.iterations >= 1
        c = og.align(scan, r.pose_vector, params)
        assert c.status == 0 and c.iterations == 1, "GPU's converged pose is not stationary for the oracle"
        ct, crr = pose_errors(r.pose_vector, c.pose)
        assert ct <= pose_m and crr <= pose_rad
        if not seen:
            return "stationary", pr.spread_m
    if pr.chaotic:
        return "chaotic", pr.spread_m
    dt, dr = pose_errors(r.pose_vector, o.pose)
    assert dt <= 2 * pr.spread_m + pose_m, f"{dt:.4f} m from the oracle, probe spread {pr.spread_m:.4f} m"
    assert dr <= 2 * pr.spread_rad + pose_rad, f"{np.rad2deg(dr):.4f} deg from the oracle, probe spread " \
                                               f"{np.rad2deg(pr.spread_rad):.4f} deg"
    return "bounded", pr.spread_m


def oracle_many(og, cases, params, threads=16):
    """Oracle alignments of (scan, pose) cases on a thread pool (the ctypes
    calls release the interpreter lock)."""
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda c: og.align(c[0], c[1], params), cases))


def check_many(results, og, cases, params, noise=1e-10, threads=16, oracle=None, hessian_rel=1e-6,
               score_rel=1e-9):
    """check_alignment over many cases in parallel; returns the outcomes."""
    if oracle is None:
        oracle = oracle_many(og, cases, params, threads)
    args = list(zip(results, cases, oracle))
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda a: check_alignment(a[0], og, a[1][0], a[1][1], params, o=a[2], noise=noise,
                                                     hessian_rel=hessian_rel, score_rel=score_rel), args))
