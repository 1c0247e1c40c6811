"""Map-independent learned initial-pose proposal, relative to a position prior.

The reference estimates an initial pose by aligning particles drawn around a
prior pose (cuda_ndt_matcher/src/initial_pose.rs:115-400): a position from
GNSS with an uncertainty of a metre or two and no usable heading.  Here the
bf16 MLP of proposal.py looks at the bird's-eye-view descriptor of the scan
(sensor frame, 64 x 64 x 8 cells of 1 m x 1 m x 2 m) and proposes 16
hypotheses (dx, dy, yaw, logit): the sensor's offset from the prior position
in the map frame and its heading.  Nothing in the input names the map, so one
network serves maps it was not trained on; what it can learn is what a scan
shows of its own heading (the procedural scenes' buildings are axis-aligned,
so the heading is determined up to quarter turns) and, from the loss, how to
spread its hypotheses over the prior's uncertainty.  The hypotheses are
refined by the Newton kernel in one batched launch, the best few run once
more when the best has not converged, and the plain particle search runs as
fallback when the best NVTL stays below a threshold.

Training data come from fresh procedural scenes: ray-cast scans at random
poses, labelled with the NDT pose converged from the true pose; every batch
turns each scan by a fresh heading and draws a fresh prior offset.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import proposal, synth
from .proposal import HYPOTHESES, ProposalNet, ScanStore, wrap

PRIOR_HALF = 1.0        # prior position error: U(-1, 1) m on x and y
OFFSET_UNIT = 4.0       # network units of the position offset: metres / 4, so that in the
                        # winner-takes-all loss a metre weighs about as much as 14 deg of heading
DESC_LEAF = 1.0         # the descriptor's input: the scan downsampled at 1 m
NVTL_THRESHOLD = 2.3    # fallback below this NVTL (the configuration's threshold)


@dataclass
class Request:
    """One initial-pose request: the scan to align, the descriptor's scan
    (both sensor frame), the true pose and the prior (x, y, z; no heading)."""
    scan: np.ndarray
    desc: np.ndarray
    true_pose: np.ndarray
    prior: np.ndarray


def scene_requests(scene: synth.Scene, seed: int, count: int, half: float, beams: int, az_steps: int,
                   align_leaf: float, prior_half: float = PRIOR_HALF, threads: int = 16) -> List[Request]:
    """Seeded sensor poses uniform over [-half, half]^2 of `scene`, yaw
    uniform, 1.8 m above ground; a ring lidar at each; the prior is the true
    position moved by U(-prior_half, prior_half) on x and y."""
    from concurrent.futures import ThreadPoolExecutor

    r = synth.rng(seed, 31)
    poses = [np.array([r.uniform(-half, half), r.uniform(-half, half), 1.8, 0.0, 0.0, r.uniform(-math.pi, math.pi)])
             for _ in range(count)]
    offs = r.uniform(-prior_half, prior_half, (count, 2))

    def one(k):
        p = poses[k]
        raw = scene.lidar(p[0], p[1], p[2], p[5], beams, az_steps=az_steps, seed=seed * 1000003 + k)
        desc = synth.voxel_downsample(raw, DESC_LEAF)
        scan = desc if align_leaf == DESC_LEAF else synth.voxel_downsample(raw, align_leaf)
        prior = np.array([p[0] + offs[k, 0], p[1] + offs[k, 1], p[2]])
        return Request(scan, desc, p, prior)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(count)))


def reference_poses(matcher, requests: Sequence[Request], chunk: int = 64) -> np.ndarray:
    """The NDT pose converged from the true pose, per request (the labels,
    and the target a search must reach)."""
    out = []
    for i in range(0, len(requests), chunk):
        part = requests[i:i + chunk]
        out.extend(r.pose_vector for r in matcher.align_parallel_scans([(q.scan, q.true_pose) for q in part]))
    return np.stack(out)


class TrainingSet:
    """Descriptor scans of several scenes in device memory with their label
    poses (converged NDT poses) and the scenes' true poses."""

    def __init__(self):
        self.store = ScanStore()
        self.refs = []
        self.true = []

    def add_scene(self, scene_seed: int, seed: int, count: int, extent: float = 200.0, n_boxes: int = 40,
                  n_poles: int = 100, map_points: int = 200_000, half: float = 60.0, beams: int = 64,
                  az_steps: int = 900, matcher_factory=None):
        """count scans of a fresh scene (its own map for the labels)."""
        scene = synth.Scene(scene_seed, extent, n_boxes, n_poles)
        reqs = scene_requests(scene, seed, count, half, beams, az_steps, DESC_LEAF)
        m = matcher_factory()
        m.set_target(scene.sample(map_points, 0.02, scene_seed))
        self.refs.append(reference_poses(m, reqs))
        m.close()
        self.true.append(np.stack([q.true_pose for q in reqs]))
        self.store.add([q.desc for q in reqs])

    def finish(self) -> "TrainingSet":
        self.store.finish()
        self.refs = np.concatenate(self.refs)
        self.true = np.concatenate(self.true)
        return self

    def __len__(self) -> int:
        return len(self.store)


def train_relative(net: ProposalNet, data: TrainingSet, epochs: int, batch: int = 256, lr: float = 1e-3,
                   seed: int = 0, prior_half: float = PRIOR_HALF) -> List[float]:
    """Adam on the winner-takes-all loss.  Every batch turns each scan by a
    fresh heading d ~ U(-pi, pi) about the sensor's z axis (the label yaw
    becomes yaw - d) and draws a fresh prior offset o ~ U(-prior_half,
    prior_half)^2 (the label offset is the reference position minus the
    prior, true + o), in OFFSET_UNIT metres; cosine learning-rate decay."""
    import torch

    m = len(data)
    n = m - m % batch
    rng = np.random.default_rng(seed)
    buf = torch.empty((batch, int(np.prod(proposal.BEV_SHAPE))), dtype=torch.int16, device="cuda:0")
    losses = []
    for ep in range(epochs):
        order = rng.permutation(m)[:n]
        tot = 0.0
        for b in range(0, n, batch):
            rows = order[b:b + batch]
            d = rng.uniform(-math.pi, math.pi, batch)
            o = rng.uniform(-prior_half, prior_half, (batch, 2))
            data.store.bev(rows, d, out=buf)
            ref, tru = data.refs[rows], data.true[rows]
            lab = np.stack([(ref[:, 0] - (tru[:, 0] + o[:, 0])) / OFFSET_UNIT,
                            (ref[:, 1] - (tru[:, 1] + o[:, 1])) / OFFSET_UNIT, wrap(ref[:, 5] - d)],
                           -1).astype(np.float32)
            step_lr = lr * (0.5 * (1 + math.cos(math.pi * (ep + b / n) / epochs)))
            tot += net.train_step_device(buf.data_ptr(), lab, step_lr)
        losses.append(tot / (n // batch))
    return losses


class DeviceRequest:
    """A request's two scans in device memory: the descriptor (row 0 of a
    zero-padded 32-row network batch) and the alignment points.  Buffers are
    reused across requests."""

    def __init__(self):
        self.desc = proposal.QueryScan(capacity=1 << 14)
        self.align = proposal.QueryScan(capacity=1 << 16)

    def load(self, req: Request):
        """Uploads both scans, forms the descriptor; returns (points pointer, n)."""
        self.desc.load(req.desc)
        if req.scan is req.desc:
            return self.desc.points.data_ptr(), req.desc.shape[0]
        return self.align.load_points(req.scan)


def hypotheses(net: ProposalNet, batch_ptr: int, k: int = HYPOTHESES) -> np.ndarray:
    """Top-k (dx, dy, yaw) of row 0 of a 32-row descriptor batch, by logit
    (dx, dy in metres)."""
    out = net.forward_device(batch_ptr, 32)[:1].reshape(1, -1, 4).astype(np.float64)
    order = np.argsort(-out[..., 3], axis=1)[:, :k]
    h = np.take_along_axis(out[..., :3], order[..., None], axis=1)[0]
    h[:, :2] *= OFFSET_UNIT
    return h


def relative_search(matcher, net: ProposalNet, req: Request, dev: DeviceRequest, k: int = HYPOTHESES,
                    polish: int = 4, nvtl_threshold: float = NVTL_THRESHOLD, fallback=None,
                    times: Optional[dict] = None, quarter_turns: int = 8):
    """Hypotheses -> one batched refinement (height from the prior, roll and
    pitch zero) -> when the best refined pose has not converged, the
    `polish` best once more -> below the NVTL threshold, the `quarter_turns`
    best hypotheses turned by 90, 180 and 270 degrees in one more launch (a
    scan of a street grid shows its heading only up to quarter turns) ->
    fallback(req) when the best NVTL is still below the threshold.  `times`
    accumulates seconds per stage.  Returns (best AlignResult,
    used_fallback)."""

    def lap(key, t0):
        if times is not None:
            times[key] = times.get(key, 0.0) + time.perf_counter() - t0
        return time.perf_counter()

    t = time.perf_counter()
    ptr, n = dev.load(req)
    t = lap("descriptor", t)
    hyp = hypotheses(net, dev.desc.batch.data_ptr(), k)
    t = lap("proposal", t)
    poses = [np.array([req.prior[0] + h[0], req.prior[1] + h[1], req.prior[2], 0.0, 0.0, h[2]]) for h in hyp]
    res = sorted(matcher.align_batch_device(ptr, n, poses), key=lambda r: -r.nvtl)
    t = lap("refine", t)
    if polish > 0 and not res[0].converged:
        again = matcher.align_batch_device(ptr, n, [r.pose_vector for r in res[:polish]])
        res = sorted(again + res, key=lambda r: -r.nvtl)
        t = lap("polish", t)
    if res[0].nvtl < nvtl_threshold and quarter_turns > 0:
        turned = [np.array([p[0], p[1], p[2], 0.0, 0.0, wrap(p[5] + q * math.pi / 2)])
                  for p in poses[:quarter_turns] for q in (1, 2, 3)]
        res = sorted(matcher.align_batch_device(ptr, n, turned) + res, key=lambda r: -r.nvtl)
        t = lap("quarter_turns", t)
    best = res[0]
    if best.nvtl < nvtl_threshold and fallback is not None:
        out = fallback(req), True
        lap("fallback", t)
        return out
    return best, False


def plain_prior_search(matcher, req: Request, particles: int, seed: int, polish: int = 0,
                       roll_pitch_deg: float = 1.0, ptr_n=None):
    """The plain particle search around the prior: translation U(+-2 m),
    yaw U(+-pi), roll/pitch N(0, roll_pitch_deg), height at the prior; all
    refined in one launch, best NVTL wins; with polish > 0 the polish best
    once more when the best has not converged."""
    r = np.random.default_rng(seed)
    poses = [np.array([req.prior[0] + r.uniform(-2, 2), req.prior[1] + r.uniform(-2, 2), req.prior[2],
                       r.normal(0, math.radians(roll_pitch_deg)), r.normal(0, math.radians(roll_pitch_deg)),
                       r.uniform(-math.pi, math.pi)]) for _ in range(particles)]

    def run(ps):
        if ptr_n is not None:
            return matcher.align_batch_device(ptr_n[0], ptr_n[1], ps)
        return matcher.align_batch(req.scan, ps)

    res = sorted(run(poses), key=lambda x: -x.nvtl)
    if polish > 0 and not res[0].converged:
        again = run([x.pose_vector for x in res[:polish]])
        res = sorted(again + res, key=lambda x: -x.nvtl)
    return res[0]


def success(pose, ref, tol_m: float = 0.02, tol_deg: float = 0.1) -> bool:
    """Within 2 cm / 0.1 deg of the reference pose (angles wrapped)."""
    pose, ref = np.asarray(pose, np.float64), np.asarray(ref, np.float64)
    return bool(np.abs(pose[:3] - ref[:3]).max() <= tol_m and
                np.rad2deg(np.abs(wrap(pose[3:] - ref[3:])).max()) <= tol_deg)


def hypothesis_quality(net: ProposalNet, requests: Sequence[Request], refs: np.ndarray, k: int = HYPOTHESES):
    """Per request, the distance of the closest raw hypothesis to the
    reference: (metres, degrees) of the hypothesis that minimises
    metres + degrees / 10."""
    dev = DeviceRequest()
    out = []
    for q, ref in zip(requests, refs):
        dev.desc.load(q.desc)
        hyp = hypotheses(net, dev.desc.batch.data_ptr(), k)
        dxy = np.hypot(q.prior[0] + hyp[:, 0] - ref[0], q.prior[1] + hyp[:, 1] - ref[1])
        dyaw = np.rad2deg(np.abs(wrap(hyp[:, 2] - ref[5])))
        j = int(np.argmin(dxy + dyaw / 10.0))
        out.append((float(dxy[j]), float(dyaw[j])))
    return np.array(out)


def heading_error_mod_quarter(net: ProposalNet, requests: Sequence[Request], refs: np.ndarray,
                              k: int = HYPOTHESES) -> np.ndarray:
    """Per request, the smallest heading error (degrees) of the k hypotheses
    modulo quarter turns: the part of the heading that a scan of the
    axis-aligned procedural scenes shows by itself."""
    dev = DeviceRequest()
    out = []
    for q, ref in zip(requests, refs):
        dev.desc.load(q.desc)
        hyp = hypotheses(net, dev.desc.batch.data_ptr(), k)
        d = np.abs(wrap(4.0 * (hyp[:, 2] - ref[5]))) / 4.0
        out.append(float(np.rad2deg(d.min())))
    return np.array(out)
