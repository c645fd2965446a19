"""Throughput of the warm-start U-Net's MFMA convolution kernels.

    python -m vflip_amd.bench_unet [--n 64] [--reps 20]

Times the forward pass (inference) and the training step (forward, MSE loss,
backward-data and weight-gradient convolutions) on an n^3 grid with HIP events
on the library's stream and prints one JSON line with the convolution FLOPs
(2 * 27 * Cin * Cout * voxels per layer, forward; backward-data and weight
gradient each as much again) and the achieved fraction of the MI355X dense BF16
MFMA peak (2.5 PFLOP/s). Elementwise work (pooling, concatenation, ReLU, the
loss) is included in the times and excluded from the FLOPs.
"""
import argparse
import json
import sys

PEAK_BF16_TFLOPS = 2500.0


def conv_flops(n):
    from .unet import LAYERS, cin_of
    total = 0
    for L in LAYERS:
        vox = (n >> L[3]) ** 3
        total += 2 * 27 * cin_of(L) * L[1] * vox
    return total


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    import torch
    from .unet import UNet
    n = a.n
    g = torch.Generator(device="cuda").manual_seed(0)
    rhs = torch.randn((n, n, n), device="cuda", generator=g)
    solid = (torch.rand((n, n, n), device="cuda", generator=g) < 0.1).float()
    liquid = 1.0 - solid
    target = torch.randn((n, n, n), device="cuda", generator=g)
    net = UNet(seed=1)
    out = torch.empty_like(rhs)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(3):
        net.forward(rhs, solid, liquid, 1.0, out)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(a.reps):
        net.forward(rhs, solid, liquid, 1.0, out)
    ev[1].record()
    for _ in range(2):
        net.zero_grad()
        net.train_step(rhs, solid, liquid, 1.0, target)
    torch.cuda.synchronize()
    ev[2].record()
    for _ in range(a.reps):
        net.train_step(rhs, solid, liquid, 1.0, target)
    ev[3].record()
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / a.reps
    trn_ms = ev[2].elapsed_time(ev[3]) / a.reps
    f = conv_flops(n)
    res = {"grid": n, "conv_gflop_forward": f / 1e9, "forward_ms": fwd_ms,
           "forward_tflops": f / (fwd_ms * 1e-3) / 1e12, "train_step_ms": trn_ms,
           "train_step_tflops": 3 * f / (trn_ms * 1e-3) / 1e12, "peak_bf16_tflops": PEAK_BF16_TFLOPS}
    res["forward_mfma_frac"] = res["forward_tflops"] / PEAK_BF16_TFLOPS
    res["train_mfma_frac"] = res["train_step_tflops"] / PEAK_BF16_TFLOPS
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
