"""The learned warm start's U-Net on the GPU against its CPU reference
(oracle/unet_ref.py: fp32 torch ops with bf16 rounding where the GPU stores
bf16, autograd gradients).

Tolerances: the GPU accumulates in fp32 on the matrix cores in another order
and rounds each layer's output gradient to bf16, so predictions agree to
relative L2 <= 2e-2 and per-layer gradients to <= 5e-2 (a transposed operand,
a wrong tap or channel offset gives O(1) errors).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sample(shape, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    nk, nj, ni = shape
    # smooth-ish divergence: random field blurred by a few box filters
    r = torch.randn(1, 1, nk, nj, ni, generator=g)
    for _ in range(2):
        r = torch.nn.functional.avg_pool3d(r, 3, stride=1, padding=1)
    rhs = (r[0, 0] / r.abs().max()).float()
    solid = (torch.rand(shape, generator=g) < 0.1).float()
    liquid = 1.0 - solid
    target = torch.randn(shape, generator=g) * 0.3
    return rhs, solid, liquid, target


def _layer_slices():
    from vflip_amd.unet import param_layout
    offs, boffs, n = param_layout()
    return [(wo, wo + cout * kw, bo, bo + cout) for (wo, cout, kw, _), bo in zip(offs, boffs)]


def _rel(a, b):
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 24, 32)])
def test_unet_forward_and_gradients_match_reference(cuda, shape):
    print(f"unet parity {shape}", flush=True)
    import torch
    from oracle import unet_ref
    from vflip_amd.unet import UNet, param_layout
    net = UNet(seed=3)
    assert net.n == param_layout()[2]
    rhs, solid, liquid, target = _sample(shape, 1)
    flat = net.params().cpu()
    loss_ref, grad_ref, pred_ref = unet_ref.loss_and_grads(flat, rhs, solid, liquid, 1.0, target)
    pred = net.forward(rhs.cuda(), solid.cuda(), liquid.cuda(), 1.0)
    torch.cuda.synchronize()
    assert _rel(pred.cpu(), pred_ref) <= 2e-2
    net.zero_grad()
    loss = net.train_step(rhs.cuda(), solid.cuda(), liquid.cuda(), 1.0, target.cuda())
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    g = net.grads().cpu()
    print("per-layer relative gradient error:",
          [round(_rel(g[w0:w1], grad_ref[w0:w1]), 4) for (w0, w1, _, _) in _layer_slices()], flush=True)
    for l, (w0, w1, b0, b1) in enumerate(_layer_slices()):
        assert _rel(g[w0:w1], grad_ref[w0:w1]) <= 5e-2, f"layer {l} weights"
        assert _rel(g[b0:b1], grad_ref[b0:b1]) <= 5e-2 or float(grad_ref[b0:b1].norm()) < 1e-8, f"layer {l} bias"


@pytest.mark.parametrize("n", [64, 128])
def test_unet_large_grid_matches_reference(cuda, n):
    """Grids whose convolutions run unsplit (64^3) and, at full resolution, with
    persistent workgroups (128^3). The reference runs in fp32 on the GPU (torch
    conv3d with TF32 off); same tolerances."""
    print(f"unet parity {n}^3", flush=True)
    import torch
    from oracle import unet_ref
    from vflip_amd.unet import UNet, param_layout
    net = UNet(seed=3)
    rhs, solid, liquid, target = _sample((n, n, n), 2)
    rhs, solid, liquid, target = rhs.cuda(), solid.cuda(), liquid.cuda(), target.cuda()
    flat = net.params().detach().clone()
    with torch.backends.cudnn.flags(enabled=True, allow_tf32=False):
        loss_ref, grad_ref, pred_ref = unet_ref.loss_and_grads(flat, rhs, solid, liquid, 1.0, target)
    pred = net.forward(rhs, solid, liquid, 1.0)
    torch.cuda.synchronize()
    assert _rel(pred, pred_ref) <= 2e-2
    net.zero_grad()
    loss = net.train_step(rhs, solid, liquid, 1.0, target)
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    g = net.grads()
    errs = [round(_rel(g[w0:w1], grad_ref[w0:w1]), 4) for (w0, w1, _, _) in _layer_slices()]
    print("per-layer relative gradient error:", errs, flush=True)
    for l, (w0, w1, b0, b1) in enumerate(_layer_slices()):
        assert _rel(g[w0:w1], grad_ref[w0:w1]) <= 5e-2, f"layer {l} weights"
        assert _rel(g[b0:b1], grad_ref[b0:b1]) <= 5e-2 or float(grad_ref[b0:b1].norm()) < 1e-8, f"layer {l} bias"


def test_unet_adam_reduces_loss(cuda):
    print("unet adam", flush=True)
    import torch
    from vflip_amd.unet import UNet
    net = UNet(seed=5)
    rhs, solid, liquid, _ = _sample((16, 16, 16), 2)
    target = rhs.clone() * 0.5  # learnable: a local map of the input
    r, s, q, t = rhs.cuda(), solid.cuda(), liquid.cuda(), target.cuda()
    losses = []
    for step in range(1, 31):
        net.zero_grad()
        losses.append(net.train_step(r, s, q, 1.0, t))
        net.adam(1e-3, step)
    torch.cuda.synchronize()
    assert losses[-1] < 0.5 * losses[0], losses[::5]
