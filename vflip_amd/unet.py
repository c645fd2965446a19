"""The learned warm start's network description and a torch-tensor view of
libvflip's U-Net entry points (vflip_unet_*).

The layer table mirrors the one compiled into the library: 4 levels with
16/32/64/128 channels, two 3x3x3 convolutions per level, 2x2x2 average pooling
down, nearest upsampling and concatenation with the skip up, and a final
convolution to one output channel. Parameters are stored per layer as a weight
matrix [cout][kw] with k = tap * cin + ci (tap = (dz+1)*9 + (dy+1)*3 + (dx+1);
kw = 27 * cin, or 28 * cin when a 16-channel input pads K with a zero tap),
followed by all biases.
"""
import ctypes

import torch

from ._lib import check, lib
from .api import _ptr, _stream

# (inputs, cout, relu, level, fan-in channels); an input is (tensor, channels, upsampled)
# tensors: layer index, "in" (network input), "P0".."P2" (pooled skips)
LAYERS = [
    ([("in", 16, 0)], 16, 1, 0, 3),
    ([(0, 16, 0)], 16, 1, 0, 16),
    ([("P0", 16, 0)], 32, 1, 1, 16),
    ([(2, 32, 0)], 32, 1, 1, 32),
    ([("P1", 32, 0)], 64, 1, 2, 32),
    ([(4, 64, 0)], 64, 1, 2, 64),
    ([("P2", 64, 0)], 128, 1, 3, 64),
    ([(6, 128, 0)], 128, 1, 3, 128),
    ([(7, 128, 1), (5, 64, 0)], 64, 1, 2, 192),
    ([(8, 64, 0)], 64, 1, 2, 64),
    ([(9, 64, 1), (3, 32, 0)], 32, 1, 1, 96),
    ([(10, 32, 0)], 32, 1, 1, 32),
    ([(11, 32, 1), (1, 16, 0)], 16, 1, 0, 48),
    ([(12, 16, 0)], 16, 1, 0, 16),
    ([(13, 16, 0)], 16, 0, 0, 16),
]
POOL_OF = {"P0": 1, "P1": 3, "P2": 5}


def cin_of(layer):
    return sum(c for _, c, _ in layer[0])


def kw_of(layer):
    ins = layer[0]
    ck = 32 if all(c % 32 == 0 for _, c, _ in ins) else 16
    return (27 if ck == 32 else 28) * cin_of(layer)


def param_layout():
    """[(w_offset, cout, kw, cin)], bias offsets, total count."""
    offs, nw = [], 0
    for L in LAYERS:
        offs.append((nw, L[1], kw_of(L), cin_of(L)))
        nw += L[1] * kw_of(L)
    boffs, nb = [], 0
    for L in LAYERS:
        boffs.append(nw + nb)
        nb += L[1]
    return offs, boffs, nw + nb


def split_params(flat):
    """Flat parameter vector -> list of (weight [cout, 27, cin], bias [cout]) (padding dropped)."""
    offs, boffs, n = param_layout()
    assert flat.numel() == n
    out = []
    for (wo, cout, kw, cin), bo in zip(offs, boffs):
        w = flat[wo:wo + cout * kw].view(cout, kw)[:, :27 * cin].reshape(cout, 27, cin)
        out.append((w, flat[bo:bo + cout]))
    return out


class UNet:
    """Owns a vflip_unet; tensors are device float32, grids (nk, nj, ni) with dims multiples of 8."""

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        check(lib().vflip_unet_create(seed, ctypes.byref(h)), "unet_create")
        self.h = h
        check(lib().vflip_unet_set_stream(self.h, _stream()), "unet_set_stream")
        self.n = int(lib().vflip_unet_num_params(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().vflip_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = torch.empty(self.n, device="cuda")
        check(lib().vflip_unet_get_params(self.h, _ptr(out)), "unet_get_params")
        return out

    def set_params(self, flat):
        flat = flat.to(device="cuda", dtype=torch.float32).contiguous()
        check(lib().vflip_unet_set_params(self.h, _ptr(flat, numel=self.n)), "unet_set_params")

    def grads(self):
        out = torch.empty(self.n, device="cuda")
        check(lib().vflip_unet_get_grads(self.h, _ptr(out)), "unet_get_grads")
        return out

    def zero_grad(self):
        check(lib().vflip_unet_zero_grad(self.h), "unet_zero_grad")

    def activation(self, tensor, shape):
        """Stored activation of the last forward as (C, z, y, x) float32: tensor 0..13 = layer
        outputs, 15..17 = pooled inputs of levels 1..3, 18 = the packed input; shape = the
        (z, y, x) grid of the tensor's level."""
        z, y, x = shape
        ch = 16 if tensor == 18 else (LAYERS[tensor][1] if tensor < len(LAYERS) else
                                      LAYERS[POOL_OF[f"P{tensor - len(LAYERS)}"]][1])
        out = torch.empty((z, y, x, ch), device="cuda")
        check(lib().vflip_unet_get_activation(self.h, tensor, _ptr(out)), "unet_get_activation")
        return out.permute(3, 0, 1, 2).contiguous()

    def forward(self, rhs, solid=None, liquid=None, in_scale=1.0, out=None):
        nk, nj, ni = rhs.shape
        out = torch.empty_like(rhs) if out is None else out
        check(lib().vflip_unet_forward(self.h, ni, nj, nk, _ptr(rhs), _ptr(solid, numel=rhs.numel()),
                                       _ptr(liquid, numel=rhs.numel()), in_scale, _ptr(out, numel=rhs.numel())),
              "unet_forward")
        return out

    def backward(self, grad_pred):
        check(lib().vflip_unet_backward(self.h, _ptr(grad_pred)), "unet_backward")

    def train_step(self, rhs, solid, liquid, in_scale, target):
        """Forward, MSE loss 0.5*mean((pred - target)^2), backward (gradients accumulate)."""
        nk, nj, ni = rhs.shape
        loss = ctypes.c_double()
        check(lib().vflip_unet_train_step(self.h, ni, nj, nk, _ptr(rhs), _ptr(solid, numel=rhs.numel()),
                                          _ptr(liquid, numel=rhs.numel()), in_scale,
                                          _ptr(target, numel=rhs.numel()), ctypes.byref(loss)), "unet_train_step")
        return loss.value

    def train_step_energy(self, rhs, solid, liquid, target, diag, ax, ay, az):
        """Forward, relative energy-norm error of the (normalised) prediction under
        the sample's operator, backward (gradients accumulate). Returns the loss."""
        nk, nj, ni = rhs.shape
        n = rhs.numel()
        loss = ctypes.c_double()
        check(lib().vflip_unet_train_step_energy(self.h, ni, nj, nk, _ptr(rhs), _ptr(solid, numel=n),
                                                 _ptr(liquid, numel=n), _ptr(target, numel=n),
                                                 *[_ptr(a, numel=n) for a in (diag, ax, ay, az)],
                                                 ctypes.byref(loss)), "unet_train_step_energy")
        return loss.value

    def adam(self, lr, step, grad_scale=1.0):
        check(lib().vflip_unet_adam(self.h, lr, step, grad_scale), "unet_adam")

    def warm_start(self, rhs, diag, factor=1, out=None):
        """Pressure guess on rhs's grid (network run on the grid coarsened by factor)."""
        nk, nj, ni = rhs.shape
        out = torch.empty_like(rhs) if out is None else out
        check(lib().vflip_unet_warm_start(self.h, ni, nj, nk, factor, _ptr(rhs), _ptr(diag, numel=rhs.numel()),
                                          _ptr(out, numel=rhs.numel())), "unet_warm_start")
        return out

    @property
    def gain(self):
        return float(lib().vflip_unet_gain(self.h))

    @gain.setter
    def gain(self, g):
        check(lib().vflip_unet_set_gain(self.h, float(g)), "unet_set_gain")

    def save(self, path):
        """Parameters (float16; the kernels use bf16 copies) and the pressure gain to an .npz."""
        import numpy as np
        np.savez_compressed(path, params=self.params().cpu().numpy().astype(np.float16),
                            gain=np.array([self.gain], np.float32))

    def load(self, path):
        import numpy as np
        with np.load(path, allow_pickle=False) as z:
            flat = torch.from_numpy(z["params"].astype(np.float32))
            gain = float(z["gain"][0])
        if flat.numel() != self.n:
            raise ValueError(f"{path}: {flat.numel()} parameters, the network has {self.n}")
        self.set_params(flat)
        self.gain = gain
