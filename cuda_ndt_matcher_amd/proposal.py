"""Learned initial-pose proposal (MLP over a bird's-eye-view of the scan).

The reference estimates an initial pose by aligning particles drawn around
a prior (cuda_ndt_matcher/src/initial_pose.rs:115-400).  Here a multilayer
perceptron looks at a 64 x 64 x 8 bird's-eye-view histogram of the scan
(1 m cells, 2 m height bins, sensor frame) and proposes H = 16 map poses
(x, y, yaw, logit).  The best proposals are refined by the Newton kernel in
one batched launch; when the best refined NVTL stays below a threshold the
plain particle search runs instead.  The network is map-specific: with only
the scan as input it can place the sensor only in the map it was trained on.

Training, inference and the descriptor run on the GPU (libndt_hip.so:
bf16 MFMA products with FP32 accumulation, FP32 Adam masters).
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib

BEV_SHAPE = (64, 64, 8)     # nx, ny, nz
BEV_CELL = 1.0              # metres, x and y
BEV_ZMIN, BEV_ZCELL = -2.0, 2.0
HYPOTHESES = 16
NET_SIZES = (64 * 64 * 8, 512, 512, 512, 4 * HYPOTHESES)


class ProposalNet:
    """The MLP: sizes[0] -> ... -> sizes[-1], GELU between layers."""

    def __init__(self, sizes: Sequence[int] = NET_SIZES, seed: int = 0):
        self._lib = _lib.load()
        self.sizes = tuple(int(s) for s in sizes)
        h = C.c_void_p()
        rc = self._lib.ndt_hip_mlp_create(np.asarray(self.sizes, np.int32), len(self.sizes), seed, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"ndt_hip_mlp_create failed ({rc})")
        self._h = h

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.ndt_hip_mlp_destroy(h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # No HIP calls during interpreter shutdown (the runtime may be gone;
        # module globals such as sys may already be cleared, hence the bound
        # default argument).
        if _finalizing():
            return
        self.close()

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._lib.ndt_hip_mlp_last_error(self._h)
            raise RuntimeError(f"mlp error {rc}: {msg.decode() if msg else ''}")

    @property
    def num_params(self) -> int:
        return int(self._lib.ndt_hip_mlp_num_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        self._check(self._lib.ndt_hip_mlp_get_params(self._h, p))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        self._check(self._lib.ndt_hip_mlp_set_params(self._h, p))

    def forward_device(self, d_x: int, batch: int) -> np.ndarray:
        """Forward of bf16 device inputs; the caller makes sure d_x is written
        (the library's stream is not ordered after other streams)."""
        out = np.zeros((batch, self.sizes[-1]), np.float32)
        self._check(self._lib.ndt_hip_mlp_forward_device(self._h, d_x, batch, out))
        return out

    def forward_backward(self, x, dout):
        x = np.ascontiguousarray(x, np.float32)
        dout = np.ascontiguousarray(dout, np.float32)
        out = np.zeros((x.shape[0], self.sizes[-1]), np.float32)
        g = np.zeros(self.num_params, np.float32)
        self._check(self._lib.ndt_hip_mlp_forward_backward(self._h, x, x.shape[0], dout, out, g))
        return out, g

    def train_step_device(self, d_x: int, labels, lr: float) -> float:
        labels = np.ascontiguousarray(labels, np.float32)
        loss = C.c_float()
        self._check(self._lib.ndt_hip_mlp_train_step_device(self._h, d_x, labels, labels.shape[0], lr,
                                                             C.byref(loss)))
        return float(loss.value)


def bev_device(d_xyz: int, d_offsets: int, rows: int, d_out: int, d_index: int = 0, d_rot: int = 0,
               shape=BEV_SHAPE, cell=BEV_CELL, zmin=BEV_ZMIN, zcell=BEV_ZCELL) -> None:
    """Bird's-eye-view descriptors (device): row j of d_out describes scan
    d_index[j] (or j when d_index is 0) of the concatenated scans, rotated
    about z by the (cos, sin) pair d_rot[j] when d_rot is not 0."""
    rc = _lib.load().ndt_hip_bev_device(d_xyz, d_offsets, d_index or None, rows, d_rot or None, shape[0], shape[1],
                                        shape[2], cell, zmin, zcell, d_out)
    if rc != 0:
        raise RuntimeError(f"ndt_hip_bev_device failed ({rc})")


class ScanStore:
    """Concatenated sensor-frame scans resident in device memory, so that
    descriptors of any rows at any rotation can be formed on the GPU.

    Only points that some rotation about z can bring into the descriptor's
    grid are kept: horizontal range within the grid's half-diagonal plus a
    cell, height within the grid's z range plus a bin either side.  The
    descriptors are the same as those of the whole scans."""

    def __init__(self, shape=BEV_SHAPE, cell=BEV_CELL, zmin=BEV_ZMIN, zcell=BEV_ZCELL):
        self._xyz, self._counts = [], []
        self.xyz = self.offsets = None
        self._r2 = (math.hypot(shape[0] / 2 * cell, shape[1] / 2 * cell) + cell) ** 2
        self._z = (zmin - zcell, zmin + (shape[2] + 1) * zcell)

    def crop(self, scan) -> np.ndarray:
        """The points of `scan` that some rotation about z can bin (host)."""
        a = np.asarray(scan, np.float32).reshape(-1, 3)
        keep = ((a[:, 0].astype(np.float64) ** 2 + a[:, 1].astype(np.float64) ** 2 <= self._r2)
                & (a[:, 2] >= self._z[0]) & (a[:, 2] < self._z[1]))
        return np.ascontiguousarray(a[keep])

    def add(self, scans: Sequence[np.ndarray]) -> None:
        import torch

        for s in scans:
            a = self.crop(s)
            self._xyz.append(torch.from_numpy(a).to("cuda:0"))
            self._counts.append(a.shape[0])

    def finish(self) -> "ScanStore":
        import torch

        total = int(np.sum(self._counts))
        if total >= 2 ** 32:
            raise ValueError(f"{total} points exceed the 32-bit offsets of the descriptor kernel")
        self.xyz = torch.cat(self._xyz).contiguous()
        # unsigned 32-bit offsets, carried in an int32 tensor of the same bits
        off = np.concatenate([[0], np.cumsum(self._counts)]).astype(np.uint32).view(np.int32)
        self.offsets = torch.from_numpy(off).to("cuda:0")
        self._xyz = []
        return self

    def __len__(self) -> int:
        return len(self._counts)

    def bev(self, rows: np.ndarray, yaw: Optional[np.ndarray] = None, out=None):
        """Descriptors of scans `rows`, each rotated by yaw[j] (radians) about
        z when yaw is given; bf16 tensor (len(rows) x 32768) on cuda:0."""
        import torch

        dev = torch.device("cuda:0")
        idx = torch.from_numpy(np.ascontiguousarray(rows, np.int32)).to(dev)
        rot = None
        if yaw is not None:
            cs = np.stack([np.cos(yaw), np.sin(yaw)], -1).astype(np.float32)
            rot = torch.from_numpy(np.ascontiguousarray(cs)).to(dev)
        if out is None:
            out = torch.empty((len(rows), int(np.prod(BEV_SHAPE))), dtype=torch.int16, device=dev)
        torch.cuda.synchronize()
        bev_device(self.xyz.data_ptr(), self.offsets.data_ptr(), len(rows), out.data_ptr(), idx.data_ptr(),
                   rot.data_ptr() if rot is not None else 0)
        return out


class QueryScan:
    """One scan at a time in device memory for the search: its descriptor in
    row 0 of a zero-padded 32-row network batch, and the points for
    align_batch_device.  One pinned upload per scan; buffers are reused."""

    def __init__(self, capacity: int = 1 << 16):
        import torch

        dev = torch.device("cuda:0")
        self.batch = torch.zeros((32, int(np.prod(BEV_SHAPE))), dtype=torch.int16, device=dev)
        self._off_h = torch.zeros(2, dtype=torch.int32).pin_memory()
        self._off = torch.zeros(2, dtype=torch.int32, device=dev)
        self._cap = 0
        self._grow(capacity)

    def _grow(self, n: int) -> None:
        import torch

        self._cap = max(n, 2 * self._cap)
        self._h = torch.empty((self._cap, 3), dtype=torch.float32).pin_memory()
        self.points = torch.empty((self._cap, 3), dtype=torch.float32, device="cuda:0")

    def load_points(self, scan: np.ndarray):
        """Uploads `scan` only (no descriptor); returns (device pointer, n)
        once the copy is complete."""
        import torch

        a = np.ascontiguousarray(np.asarray(scan, np.float32).reshape(-1, 3))
        n = a.shape[0]
        if n > self._cap:
            self._grow(n)
        self._h[:n].numpy()[...] = a
        self.points[:n].copy_(self._h[:n], non_blocking=True)
        torch.cuda.synchronize()
        return self.points.data_ptr(), n

    def load(self, scan: np.ndarray):
        """Uploads `scan` and forms its descriptor; returns (device pointer, n)."""
        import torch

        a = np.ascontiguousarray(np.asarray(scan, np.float32).reshape(-1, 3))
        n = a.shape[0]
        if n > self._cap:
            self._grow(n)
        self._h[:n].numpy()[...] = a
        self.points[:n].copy_(self._h[:n], non_blocking=True)
        self._off_h[1] = n
        self._off.copy_(self._off_h, non_blocking=True)
        # The descriptor kernel runs on another stream than torch's copies.
        torch.cuda.synchronize()
        bev_device(self.points.data_ptr(), self._off.data_ptr(), 1, self.batch.data_ptr())
        return self.points.data_ptr(), n


def bev_batch(scans: Sequence[np.ndarray], yaw: Optional[np.ndarray] = None):
    """Descriptors of host scans (rotated by yaw[j] about z when given), as a
    bf16 torch tensor (batch x 32768) on cuda:0."""
    store = ScanStore()
    store.add(scans)
    store.finish()
    return store.bev(np.arange(len(scans)), yaw)


# ---------------------------------------------------------------------------
# Labels: map poses normalised to the network's output range.
@dataclass
class Frame:
    """Output units of the network: x, y as (x - cx) / unit, (y - cy) / unit;
    yaw in radians.  A unit of a few metres weighs a metre of position about
    as much as ten-odd degrees of heading in the loss."""
    cx: float = 0.0
    cy: float = 0.0
    unit: float = 4.0

    def encode(self, x, y, yaw):
        return np.stack([(np.asarray(x) - self.cx) / self.unit, (np.asarray(y) - self.cy) / self.unit,
                         np.asarray(yaw)], -1).astype(np.float32)

    def decode(self, out: np.ndarray):
        """(B, 4H) network outputs -> (B, H, 4) of x, y, yaw, logit in map units."""
        o = out.reshape(out.shape[0], -1, 4).astype(np.float64)
        x = self.cx + o[..., 0] * self.unit
        y = self.cy + o[..., 1] * self.unit
        return np.stack([x, y, o[..., 2], o[..., 3]], -1)


# ---------------------------------------------------------------------------
# Data, training and the proposal-then-refine search.
@dataclass
class Sample:
    raw: np.ndarray        # lidar returns, sensor frame
    scan: np.ndarray       # the same, downsampled at 1 m (alignment and descriptor input)
    true_pose: np.ndarray  # [x, y, z, roll, pitch, yaw]


def make_samples(scene_seed: int, seed: int, count: int, half_extent: float, beams: int = 32, az_steps: int = 900,
                 threads: int = 16) -> List[Sample]:
    """Seeded sensor poses uniform over [-half_extent, half_extent]^2 of the
    small workload's scene (the one its map is sampled from: scene_seed is
    the map's seed), yaw uniform in [-pi, pi), 1.8 m above ground; a ring
    lidar per pose.  `seed` draws the poses and the range noise.  The ray
    casts run on `threads` host threads (the ray caster releases the
    interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    from . import synth

    extent, nb, npl = synth.WORKLOADS["small"][:3]
    scene = synth.Scene(scene_seed, extent, nb, npl)
    r = synth.rng(seed, 21)
    poses = [np.array([r.uniform(-half_extent, half_extent), r.uniform(-half_extent, half_extent), 1.8, 0.0, 0.0,
                       r.uniform(-math.pi, math.pi)]) for _ in range(count)]

    def one(k):
        p = poses[k]
        raw = scene.lidar(p[0], p[1], p[2], p[5], beams, az_steps=az_steps, seed=seed * 1000003 + k)
        return Sample(raw, synth.voxel_downsample(raw, 1.0), p)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(count)))


def refine_labels(matcher, samples: Sequence[Sample], chunk: int = 200) -> np.ndarray:
    """Converged NDT poses from the true poses (the training labels): the
    samples' downsampled scans aligned in batched launches."""
    out = []
    for i in range(0, len(samples), chunk):
        part = samples[i:i + chunk]
        res = matcher.align_parallel_scans([(s.scan, s.true_pose) for s in part])
        out.extend(r.pose_vector for r in res)
    return np.stack(out)


def wrap(a):
    return (np.asarray(a) + math.pi) % (2 * math.pi) - math.pi


def train(net: ProposalNet, store: ScanStore, poses: np.ndarray, frame: Frame, epochs: int, batch: int = 256,
          lr: float = 1e-3, seed: int = 0, rotate: bool = True) -> List[float]:
    """Adam on the winner-takes-all loss over the scans of `store` with map
    poses `poses` (N x 6).  With `rotate`, every batch turns each scan by a
    fresh heading d ~ U(-pi, pi) about the sensor's z axis and labels it
    with yaw - d (the same place seen at another heading), so each epoch
    shows every place at new headings.  Cosine learning-rate decay."""
    import torch

    m = len(store)
    n = m - m % batch
    rng = np.random.default_rng(seed)
    buf = torch.empty((batch, int(np.prod(BEV_SHAPE))), dtype=torch.int16, device="cuda:0")
    losses = []
    for ep in range(epochs):
        order = rng.permutation(m)[:n]
        tot = 0.0
        for b in range(0, n, batch):
            rows = order[b:b + batch]
            p = poses[rows]
            d = rng.uniform(-math.pi, math.pi, batch) if rotate else None
            store.bev(rows, d, out=buf)
            yaw = p[:, 5] if d is None else wrap(p[:, 5] - d)
            step_lr = lr * (0.5 * (1 + math.cos(math.pi * (ep + b / n) / epochs)))
            tot += net.train_step_device(buf.data_ptr(), frame.encode(p[:, 0], p[:, 1], yaw), step_lr)
        losses.append(tot / (n // batch))
    return losses


def propose(net: ProposalNet, frame: Frame, bevs, k: int) -> np.ndarray:
    """Top-k hypotheses (x, y, yaw) per row of bevs (batch padded to 32)."""
    import torch

    b = bevs.shape[0]
    pad = (-b) % 32
    x = bevs if pad == 0 else torch.cat([bevs, bevs[:1].expand(pad, -1)]).contiguous()
    # The network runs on the library's own stream, which is not ordered after
    # torch's: the padded batch must be complete before the launch.
    torch.cuda.synchronize()
    out = net.forward_device(x.data_ptr(), x.shape[0])[:b]
    hyp = frame.decode(out)
    order = np.argsort(-hyp[..., 3], axis=1)[:, :k]
    return np.take_along_axis(hyp[..., :3], order[..., None], axis=1)


def learned_search(matcher, net: ProposalNet, frame: Frame, sample: Sample, bev_row, k: int = HYPOTHESES,
                   nvtl_threshold: float = 2.3, polish: int = 4, fallback=None, times: Optional[dict] = None,
                   query: Optional[QueryScan] = None):
    """Proposals -> one batched refinement of the top k.  When the best
    refined pose (by NVTL) has not converged, the `polish` best refined poses
    run once more from where they stopped (one more batched launch: a
    hypothesis a metre away can spend its iteration budget on 0.1-long steps
    before it converges).  The plain search (fallback(sample)) runs when the
    best NVTL is still below the threshold.  Height comes from the prior
    (the sample's pose), roll and pitch start at zero.
    With a QueryScan the scan is uploaded once, its descriptor formed on the
    GPU (bev_row is ignored) and the refinements read the device copy.
    `times`, if given, accumulates seconds per stage (descriptor, proposal,
    refine, polish, fallback).
    Returns (best AlignResult, used_fallback)."""
    import time

    def lap(key, t0):
        if times is not None:
            times[key] = times.get(key, 0.0) + time.perf_counter() - t0
        return time.perf_counter()

    t = time.perf_counter()
    if query is not None:
        ptr, n = query.load(sample.scan)
        t = lap("descriptor", t)
        out = net.forward_device(query.batch.data_ptr(), query.batch.shape[0])[:1]
        h = frame.decode(out)
        order = np.argsort(-h[..., 3], axis=1)[:, :k]
        hyp = np.take_along_axis(h[..., :3], order[..., None], axis=1)[0]

        def refine(poses):
            return matcher.align_batch_device(ptr, n, poses)
    else:
        hyp = propose(net, frame, bev_row, k)[0]

        def refine(poses):
            return matcher.align_batch(sample.scan, poses)
    t = lap("proposal", t)
    z = sample.true_pose[2]
    poses = [np.array([h[0], h[1], z, 0.0, 0.0, h[2]]) for h in hyp]
    res = sorted(refine(poses), key=lambda r: -r.nvtl)
    t = lap("refine", t)
    if polish > 0 and not res[0].converged:
        again = refine([r.pose_vector for r in res[:polish]])
        res = sorted(again + res, key=lambda r: -r.nvtl)
        t = lap("polish", t)
    best = res[0]
    if best.nvtl < nvtl_threshold and fallback is not None:
        out = fallback(sample), True
        lap("fallback", t)
        return out
    return best, False


def plain_search(matcher, sample: Sample, particles: int, seed: int, polish: int = 0):
    """Particles around the true pose: translation U(+-2 m), yaw U(+-pi),
    roll/pitch N(0, 1 deg); all refined in one launch; best NVTL wins.  With
    polish > 0 the polish best run once more when the best has not converged
    (the learned search's second launch)."""
    r = np.random.default_rng(seed)
    p0 = sample.true_pose
    poses = [p0 + np.array([r.uniform(-2, 2), r.uniform(-2, 2), 0.0, r.normal(0, math.radians(1)),
                            r.normal(0, math.radians(1)), r.uniform(-math.pi, math.pi)]) for _ in range(particles)]
    res = sorted(matcher.align_batch(sample.scan, poses), key=lambda x: -x.nvtl)
    if polish > 0 and not res[0].converged:
        again = matcher.align_batch(sample.scan, [x.pose_vector for x in res[:polish]])
        res = sorted(again + res, key=lambda x: -x.nvtl)
    return res[0]
