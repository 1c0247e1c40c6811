"""CPU reference of the proposal network (test infrastructure only).

numpy restatement of the forward and backward passes computed by the MFMA
kernels in cuda_ndt_matcher_amd/csrc/mlp.hip, used by the GPU tests to check
the kernels' outputs and gradients.  `mimic_bf16=True` rounds the operands
of every product to bf16 exactly where the kernels do (inputs, weights,
hidden activations, and the output-gradient factors of the backward
products), so the comparison isolates accumulation order; with
`mimic_bf16=False` it is plain FP32 (the mathematical reference).
Also: the winner-takes-all proposal loss and the bird's-eye-view histogram.
"""
import math

import numpy as np


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def _erf(x):
    v = np.vectorize(math.erf)
    return v(x.astype(np.float64)).astype(np.float32)


def gelu(z):
    return (0.5 * z * (1.0 + _erf(z / np.sqrt(2.0)))).astype(np.float32)


def gelu_grad(z):
    cdf = 0.5 * (1.0 + _erf(z / np.sqrt(2.0)))
    pdf = np.exp(-0.5 * z.astype(np.float64) ** 2) / np.sqrt(2.0 * np.pi)
    return (cdf + z * pdf).astype(np.float32)


def unpack(params, sizes):
    """Per layer (W^T (out x in), b) from the flat parameter vector."""
    out, o = [], 0
    for k, n in zip(sizes[:-1], sizes[1:]):
        w = params[o:o + n * k].reshape(n, k)
        o += n * k
        b = params[o:o + n]
        o += n
        out.append((w, b))
    return out


def forward_backward(params, sizes, x, dout, mimic_bf16=True):
    r = bf16 if mimic_bf16 else (lambda a: np.asarray(a, np.float32))
    layers = unpack(np.asarray(params, np.float32), sizes)
    acts, zs = [r(x)], []
    for i, (w, b) in enumerate(layers):
        z = (acts[-1].astype(np.float64) @ r(w).T.astype(np.float64) + b).astype(np.float32)
        zs.append(z)
        if i + 1 < len(layers):
            acts.append(r(gelu(z)))
    out = zs[-1]
    grads = []
    dy = np.asarray(dout, np.float32)
    for i in range(len(layers) - 1, -1, -1):
        w, b = layers[i]
        dz = dy if i + 1 == len(layers) else (dy * gelu_grad(zs[i])).astype(np.float32)
        gw = (r(dz).T.astype(np.float64) @ acts[i].astype(np.float64)).astype(np.float32)
        gb = dz.sum(0, dtype=np.float64).astype(np.float32)
        grads.append((gw, gb))
        if i > 0:
            dy = (r(dz).astype(np.float64) @ r(w).astype(np.float64)).astype(np.float32)
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in reversed(grads)])
    return out, flat


def wta_loss(out, labels, eps=0.05):
    """Mean winner-takes-all loss and its gradient (see wta_loss in mlp.hip)."""
    out = np.asarray(out, np.float64)
    B = out.shape[0]
    H = out.shape[1] // 4
    o = out.reshape(B, H, 4)
    lx, ly, lt = labels[:, 0:1], labels[:, 1:2], labels[:, 2:3]
    e = (o[..., 0] - lx) ** 2 + (o[..., 1] - ly) ** 2 + 2 * (1 - np.cos(o[..., 2] - lt))
    best = e.argmin(1)
    w = np.full((B, H), eps / H)
    w[np.arange(B), best] += 1.0
    logits = o[..., 3]
    m = logits.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(logits - m).sum(1))
    loss = (w * e).sum(1) - (logits[np.arange(B), best] - lse)
    p = np.exp(logits - lse[:, None])
    g = np.zeros_like(o)
    g[..., 0] = w * 2 * (o[..., 0] - lx) / B
    g[..., 1] = w * 2 * (o[..., 1] - ly) / B
    g[..., 2] = w * 2 * np.sin(o[..., 2] - lt) / B
    onehot = np.zeros((B, H))
    onehot[np.arange(B), best] = 1
    g[..., 3] = (p - onehot) / B
    return loss.mean(), g.reshape(B, 4 * H)


def bev(points, nx=64, ny=64, nz=8, cell=1.0, zmin=-2.0, zcell=2.0, rot=None):
    """log(1 + count) / 4 per (iy, ix, iz) cell, flattened like the kernel;
    rot = (c, s) rotates x/y first, every operation rounded to f32."""
    p = np.asarray(points, np.float32)
    x, y = p[:, 0], p[:, 1]
    if rot is not None:
        c, s = np.float32(rot[0]), np.float32(rot[1])
        x, y = c * x - s * y, s * x + c * y
    ix = np.floor(x / np.float32(cell)).astype(np.int64) + nx // 2
    iy = np.floor(y / np.float32(cell)).astype(np.int64) + ny // 2
    iz = np.floor((p[:, 2] - np.float32(zmin)) / np.float32(zcell)).astype(np.int64)
    ok = (ix >= 0) & (ix < nx) & (iy >= 0) & (iy < ny) & (iz >= 0) & (iz < nz)
    c = np.zeros(nx * ny * nz)
    np.add.at(c, ((iy[ok] * nx + ix[ok]) * nz + iz[ok]), 1)
    return bf16(np.log1p(c).astype(np.float32) * 0.25)
