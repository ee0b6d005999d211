"""CPU reference of the warm-start network -- TEST INFRASTRUCTURE ONLY.

Only tests/ (and smoke/bench checkers) may import this module; the product
path (libaphid_gpu.so, aphid_unet_*) never calls it.

The network has no counterpart in the reference code (which always starts
QMR from x0 = 0, src/aphid_qmr.f90:230-233).  This file restates the
3-level 3-D U-Net of aphid_git_amd/csrc/unet.hip with torch.nn.functional on
the CPU in float64.  It is an independent implementation of the same
function: torch's conv3d, avg_pool3d, nearest interpolation and autograd.
The architecture table below is the only thing the two share.

The GPU stores some tensors in bf16, and this reference rounds at the same
points:
  - forward: the input features, the weights, every stored activation (after
    the ReLU) and every pooled tensor are rounded to bf16.  The rounding
    passes gradients unchanged (straight-through), as the GPU's backward
    does;
  - backward: the gradient at each convolution's pre-activation is rounded
    to bf16, because the GPU feeds it to the matrix cores as bf16;
  - the loss is the mean squared error against the bf16-rounded target.
Everything else is exact in float64.  The GPU accumulates in FP32, so
agreement is to FP32 accumulation order plus rare one-ulp bf16 flips.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np
import torch
import torch.nn.functional as F

# (kernel size, cin, cout): the layer order of aphid_git_amd/csrc/unet.hip
LAYERS: List[Tuple[int, int, int]] = [
    (3, 3, 16),   # 0 e1a
    (3, 16, 16),  # 1 e1b -> skip 1
    (3, 16, 32),  # 2 e2a
    (3, 32, 32),  # 3 e2b -> skip 2
    (3, 32, 64),  # 4 b1
    (3, 64, 64),  # 5 b2
    (1, 64, 32),  # 6 r3, then up x2
    (3, 64, 32),  # 7 d2a on [up | skip 2]
    (3, 32, 32),  # 8 d2b
    (1, 32, 16),  # 9 r2, then up x2
    (3, 32, 16),  # 10 d1a on [up | skip 1]
    (3, 16, 16),  # 11 d1b
    (1, 16, 8),   # 12 out (no ReLU)
]


def n_params() -> int:
    return sum(co * ci * k ** 3 + co for k, ci, co in LAYERS)


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to bf16, returned as float32 (numpy)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


class _RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g


class _GradBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.float32).to(torch.bfloat16).to(g.dtype)


def rb(x):
    return _RoundBF16.apply(x)


def gb(x):
    return _GradBF16.apply(x)


def split_params(flat, requires_grad=False):
    """Flat FP32 parameters -> [(w [cout,cin,k,k,k], b [cout])] float64."""
    flat = torch.as_tensor(np.asarray(flat, np.float64))
    out, off = [], 0
    for k, ci, co in LAYERS:
        nw = co * ci * k ** 3
        w = flat[off:off + nw].reshape(co, ci, k, k, k).clone()
        off += nw
        b = flat[off:off + co].clone()
        off += co
        if requires_grad:
            w.requires_grad_(True)
            b.requires_grad_(True)
        out.append((w, b))
    return out


def _conv(x, wb, k, relu=True):
    w, b = wb
    y = gb(F.conv3d(x, rb(w), bias=b, padding=k // 2))
    return rb(F.relu(y)) if relu else y


def forward(params, feat: np.ndarray, cells: Tuple[int, int, int], keep=None):
    """feat [V][3] (z fastest) -> out [1][8][X][Y][Z] (float64 tensor)."""
    X, Y, Z = cells
    x = torch.as_tensor(np.asarray(feat, np.float64)).reshape(X, Y, Z, 3).permute(3, 0, 1, 2)[None]
    x = rb(x.contiguous())
    P = params
    a1 = _conv(x, P[0], 3)
    s1 = _conv(a1, P[1], 3)
    p1 = rb(F.avg_pool3d(s1, 2))
    a2 = _conv(p1, P[2], 3)
    s2 = _conv(a2, P[3], 3)
    p2 = rb(F.avg_pool3d(s2, 2))
    a3 = _conv(p2, P[4], 3)
    b3 = _conv(a3, P[5], 3)
    r3 = _conv(b3, P[6], 1)
    c2 = torch.cat([F.interpolate(r3, scale_factor=2, mode="nearest"), s2], dim=1)
    d2a = _conv(c2, P[7], 3)
    d2b = _conv(d2a, P[8], 3)
    r2 = _conv(d2b, P[9], 1)
    c1 = torch.cat([F.interpolate(r2, scale_factor=2, mode="nearest"), s1], dim=1)
    d1a = _conv(c1, P[10], 3)
    d1b = _conv(d1a, P[11], 3)
    out = _conv(d1b, P[12], 1, relu=False)
    if keep is not None:
        keep.update(dict(a1=a1, cat1=c1, p1=p1, a2=a2, cat2=c2, p2=p2, a3=a3, b3=b3, r3=r3,
                         d2a=d2a, d2b=d2b, r2=r2, d1a=d1a, d1b=d1b, out=out))
    return out


def to_voxels(t) -> np.ndarray:
    """[1][C][X][Y][Z] tensor -> [V][C] numpy (z fastest)."""
    t = t.detach()[0]
    return t.permute(1, 2, 3, 0).reshape(-1, t.shape[0]).numpy()


def loss_grad(flat, feat, target, cells):
    """Mean squared error against bf16(target) and its gradient (flat)."""
    P = split_params(flat, requires_grad=True)
    out = forward(P, feat, cells)
    t = torch.as_tensor(bf16_round(np.asarray(target, np.float32)).astype(np.float64))
    loss = ((to_voxels_t(out) - t) ** 2).mean()
    loss.backward()
    g = torch.cat([torch.cat([w.grad.reshape(-1), b.grad]) for w, b in P])
    return float(loss), g.numpy()


def backward_from(flat, feat, cells, dout):
    """Parameter gradient of sum(out * dout) for a given output gradient
    dout [V][8]: the network's backward pass from an arbitrary d loss /
    d output (as a loss computed outside the network produces)."""
    P = split_params(flat, requires_grad=True)
    out = forward(P, feat, cells)
    d = torch.as_tensor(np.asarray(dout, np.float64))
    (to_voxels_t(out) * d).sum().backward()
    return torch.cat([torch.cat([w.grad.reshape(-1), b.grad]) for w, b in P]).numpy()


def to_voxels_t(t):
    t = t[0]
    return t.permute(1, 2, 3, 0).reshape(-1, t.shape[0])
