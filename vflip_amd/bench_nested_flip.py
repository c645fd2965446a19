"""Throughput of the nested FLIP coupling calls on one GPU.

Workload: a 128^3 coarse grid (h = 1/128) with 8 jittered particles per cell
(16.7M, the FLIP smoke fill), one subdomain over [0.3, 0.7]^3 at 2x (102^3 fine
cells), a sphere of radius 0.2 at the centre as the solid on both grids, seeded
random MAC velocities on both, solid velocity (0.2, -0.1, 0.05). Per iteration,
in the order of the reference's step (S/Smoke_solver3D.cpp:330-411):

* seed_ghost over the subdomain, then its binning (clamp 1) over real + ghost
  particles and the PSE viscosity of the real ones (diffuse);
* advect_FLIP of all coarse particles through the blended velocity;
* the halo copy of the coarse particles into the subdomain (is_halo, dx = h/2).

Each call is timed with events on the stream it runs on (the selections read
their totals back, so they include one synchronisation each).

    python -m vflip_amd.bench_nested_flip [--n 128] [--iters 10]
"""
import argparse
import json
import sys

import numpy as np
import torch

from . import api, scenes


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--iters", type=int, default=10)
    a = ap.parse_args(argv)
    n = a.n
    coarse, sub = scenes.nested_flip_case(n=n, box=(0.3, 0.7), seed=1, solid_vel=(0.2, -0.1, 0.05))
    dev = lambda d: {k: (torch.from_numpy(v).cuda() if isinstance(v, np.ndarray) else v) for k, v in d.items()}
    C, S = dev(coarse), dev(sub)
    m = sub["u"].shape[0]
    h = 1.0 / n
    g = torch.Generator(device="cuda").manual_seed(0)
    ar = torch.arange(n, device="cuda", dtype=torch.float32)
    k, j, i = torch.meshgrid(ar, ar, ar, indexing="ij")
    centre = torch.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], 1).repeat_interleave(8, 0)
    pos0 = ((centre + torch.rand(centre.shape, generator=g, device="cuda") - 0.5) * h).contiguous()
    vel0 = (torch.rand(pos0.shape, generator=g, device="cuda") * 2 - 1).contiguous()
    del centre, k, j, i
    Pc = api.Particles(n, n, n, h)
    Ps = api.Particles(m, m, m, float(sub["h"]))
    dx = float(np.float32(0.5) * np.float32(h))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
    times = {k: [] for k in ("seed_ghost", "hash_sub", "pse_diffuse", "advect_nested", "halo_select")}
    counts = {}
    sp, sv = api.halo_select(Pc, S, dx, pos0, vel0)
    for it in range(a.iters + 2):
        pos = pos0.clone()
        ev[0].record()
        gp, gv = api.seed_ghost(Ps, C, S, it)
        ev[1].record()
        allp = torch.cat([sp, gp]).contiguous()
        allv = torch.cat([sv, gv]).contiguous()
        ev[2].record()
        _, start, order, _ = Ps.hash(allp, clamp=True, bmin=sub["bmin"], count=False)
        ev[3].record()
        newv = api.pse_diffuse(C, S, 0.01, 0.05, sp.shape[0], allp, allv, start, order)
        ev[4].record()
        api.advect_particles_nested(C, S, 0.01, pos)
        ev[5].record()
        hp, hv = api.halo_select(Pc, S, dx, pos, vel0)
        ev[6].record()
        torch.cuda.synchronize()
        if it >= 2:
            times["seed_ghost"].append(ev[0].elapsed_time(ev[1]))
            times["hash_sub"].append(ev[2].elapsed_time(ev[3]))
            times["pse_diffuse"].append(ev[3].elapsed_time(ev[4]))
            times["advect_nested"].append(ev[4].elapsed_time(ev[5]))
            times["halo_select"].append(ev[5].elapsed_time(ev[6]))
        counts = {"coarse_particles": int(pos.shape[0]), "subdomain_particles": int(sp.shape[0]),
                  "ghosts": int(gp.shape[0]), "halo_after_advection": int(hp.shape[0]),
                  "diffused": int((newv != sv).any(1).sum().item())}
        print(f"iter {it}: " + ", ".join(f"{k} {v[-1]:.3f} ms" for k, v in times.items() if v), file=sys.stderr,
              flush=True)
    out = {"workload": f"nested FLIP coupling: {n}^3 coarse, 8 particles/cell, subdomain [0.3,0.7]^3 at 2x "
                       f"({m}^3 cells), sphere r=0.2",
           **counts, **{k + "_ms": float(np.mean(v)) for k, v in times.items()}}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
