"""Throughput of the FLIP particle rebuild (reorder_FLIP) and the binning before it.

Workload: the reference's FLIP smoke particle set on an n^3 grid -- 8 particles
per cell over the whole domain (SmokeSolver3D seeds 8 per cell and rebuilds
after every advection), jittered over 1.6 cells so cell counts vary (0..~20),
a closed box with a sphere of radius 0.15 at the centre as the solid, random
MAC velocities. Three cases:

* first: particles in random order (the first rebuild after seeding);
* steady: the output of one rebuild (cell-sorted), every particle moved by up
  to 0.3 h per axis -- a pessimistic step: about 39% of the particles change
  cell, and the y/z movers are gathered from far away in memory;
* still: the output of one rebuild, not moved (the gathers are contiguous).

Each timed iteration is vflip_hash (binning) then vflip_reorder (count, scan,
readback of the total, emit), timed with events on the stream both run on.
Algorithmic bytes of the rebuild (the least any rebuild moves): every binned
particle's order entry, position and velocity read once (28 B), every output
particle written once (24 B), the cell starts read once (4 B per cell).

    python -m vflip_amd.bench_reorder [--n 128] [--iters 20]
"""
import argparse
import json
import sys

import numpy as np
import torch

from . import api, scenes


def make_particles(n, seed=0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    h = 1.0 / n
    ar = torch.arange(n, device="cuda", dtype=torch.float32)
    k, j, i = torch.meshgrid(ar, ar, ar, indexing="ij")
    centre = torch.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], 1).repeat_interleave(8, 0)
    jit = (torch.rand(centre.shape, generator=g, device="cuda") - 0.5) * 1.6
    pos = ((centre + jit) * h).contiguous()
    perm = torch.randperm(pos.shape[0], generator=g, device="cuda")
    pos = pos[perm].contiguous()
    vel = (torch.rand(pos.shape, generator=g, device="cuda") * 2 - 1).contiguous()
    return pos, vel


def run_case(P, pos, vel, grids, phi, iters, seed):
    out = None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_hash, t_reo = [], []
    res = None
    for it in range(iters + 2):
        ev[0].record()
        _, start, order, _ = P.hash(pos, count=False)
        ev[1].record()
        res = P.reorder(pos, vel, start, order, *grids, phi, seed, out=out)
        ev[2].record()
        torch.cuda.synchronize()
        out = (res[0].new_empty((res[0].shape[0] + res[0].shape[0] // 8, 3)),
               res[1].new_empty((res[0].shape[0] + res[0].shape[0] // 8, 3))) if out is None else out
        if it >= 2:
            t_hash.append(ev[0].elapsed_time(ev[1]))
            t_reo.append(ev[1].elapsed_time(ev[2]))
    nbin = int(start[-1].item())
    nout = res[0].shape[0]
    return res, nbin, nout, float(np.mean(t_hash)), float(np.mean(t_reo))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--iters", type=int, default=20)
    a = ap.parse_args(argv)
    n = a.n
    h = 1.0 / n
    P = api.Particles(n, n, n, h)
    grids = [torch.from_numpy(g).cuda() for g in scenes.random_mac(n, 5, scale=1.0)]
    phi = torch.from_numpy(scenes.solid_phi(n, h, [((0.5, 0.5, 0.5), 0.15)], open_top=False)).cuda()
    pos, vel = make_particles(n)
    ncell = n ** 3
    results = {}
    first_out = None
    for case in ("first", "steady", "still"):
        (op, ov), nbin, nout, th, tr = run_case(P, pos, vel, grids, phi, a.iters, seed=1)
        byts = nbin * 28 + nout * 24 + ncell * 4
        results[case] = {"particles_in": int(pos.shape[0]), "binned": nbin, "particles_out": nout,
                         "hash_ms": th, "reorder_ms": tr, "particles_per_s": nbin / (tr * 1e-3),
                         "reorder_algorithmic_GBps": byts / (tr * 1e-3) / 1e9}
        print(f"{case}: hash {th:.3f} ms, reorder {tr:.3f} ms, {nbin} -> {nout}", file=sys.stderr, flush=True)
        if case == "first":
            first_out = (op.clone(), ov.clone())
            g = torch.Generator(device="cuda").manual_seed(7)
            pos = (op + (torch.rand(op.shape, generator=g, device="cuda") - 0.5) * (0.6 * h)).contiguous()
            vel = ov.clone()
        elif case == "steady":
            pos, vel = first_out
    print(json.dumps({"workload": f"FLIP smoke particle rebuild, {n}^3 cells, 8 particles/cell jittered over 1.6 "
                      "cells, closed box + sphere r=0.15", **results}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
