"""CPU reference of the learned warm start's U-Net (fp32 PyTorch on the CPU).

TEST INFRASTRUCTURE ONLY: the checker of libvflip's vflip_unet_* kernels, used
by tests/. It restates the network of vflip_amd/unet.py with plain torch ops
(conv3d with padding 1, 2x2x2 average pooling, nearest upsampling, channel
concatenation) and rounds to bf16 where the GPU stores bf16: the packed input,
the weights, every hidden activation and every pooled tensor. Gradients come
from torch autograd in fp32, with each convolution's output gradient rounded to
bf16 where the GPU rounds it (optional), so the remaining differences are the
fp32 accumulation orders.

There is no counterpart in the reference repository: the learned part is this
build's own design.
"""
import torch
import torch.nn.functional as F

from vflip_amd.unet import LAYERS, POOL_OF, split_params


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


class _GradBF16(torch.autograd.Function):
    """Identity forward; backward rounds the incoming gradient to bf16 (the GPU
    stores each convolution's output gradient, after the ReLU mask, as bf16)."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return _bf(g)


def _forced(a, key, forced):
    """a with the value forced[key] (the GPU's stored tensor, (C, z, y, x)) and a's gradient
    path: forced + (a - a.detach()), so the backward runs on the same forward values."""
    if forced is None or key not in forced:
        return a
    return forced[key].unsqueeze(0) + (a - a.detach())


def forward(flat, rhs, solid=None, liquid=None, in_scale=1.0, grad_bf16=True, forced=None):
    """flat: fp32 parameter vector (the library's layout, requires_grad allowed);
    rhs/solid/liquid: (nk, nj, ni) fp32 tensors (CPU, or all on one GPU for large
    grids). Returns the prediction (nk, nj, ni).
    grad_bf16: round each convolution's output gradient to bf16 in the backward
    pass, where the GPU does. forced: optional {"in", "P0".."P2", layer index:
    (C, z, y, x) tensor} of stored forward values to use in place of the ones
    recomputed here (the gradients still flow through this forward)."""
    layers = split_params(flat)
    nk, nj, ni = rhs.shape
    x = torch.zeros(1, 16, nk, nj, ni, device=rhs.device)
    x[0, 0] = _bf(rhs * in_scale)
    x[0, 1] = _bf(solid) if solid is not None else 0.0
    x[0, 2] = _bf(liquid) if liquid is not None else 1.0
    acts = {"in": _forced(x, "in", forced)}
    pred = None
    for l, (ins, cout, relu, level, _) in enumerate(LAYERS):
        for p, src in POOL_OF.items():
            if src == l - 1 and p not in acts and l in (2, 4, 6):
                acts[p] = _forced(_bf(F.avg_pool3d(acts[src], 2)), p, forced)
        parts = []
        for (t, c, up) in ins:
            a = acts[t]
            if up:
                a = F.interpolate(a, scale_factor=2, mode="nearest")
            parts.append(a)
        xin = torch.cat(parts, dim=1) if len(parts) > 1 else parts[0]
        w, b = layers[l]
        cin = w.shape[2]
        W = _bf(w).reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
        y = F.conv3d(xin, W, b, padding=1)
        if grad_bf16:
            y = _GradBF16.apply(y)
        if l == len(LAYERS) - 1:
            pred = y[0, 0]
        else:
            acts[l] = _forced(_bf(F.relu(y) if relu else y), l, forced)
    return pred


def loss_and_grads(flat, rhs, solid, liquid, in_scale, target):
    """0.5*mean((pred - target)^2) and its gradient with respect to flat."""
    p = flat.detach().clone().requires_grad_(True)
    pred = forward(p, rhs, solid, liquid, in_scale)
    loss = 0.5 * ((pred - target) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.detach(), pred.detach()
