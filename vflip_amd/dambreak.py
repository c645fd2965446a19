"""FLIP liquid dam break on the GPU: the transfer kernels of the smoke/FLIP path
(hash_FLIP, particle_to_grid, particle_from_grid, advect_FLIP) with a
free-surface pressure projection.

Workload: an n^3 closed box (walls one cell in, as the smoke scenes), a block of
liquid [0,0.4] x [0,0.6] x [0,1] filled with `ppc` jittered particles per cell
(seeded), gravity -9.8 along y, FLIP/PIC ratio 0.95, CFL-limited time step
dt = min(max_dt, cfl h / max|v_particle|). One step:

1. bin the particles (vflip_hash) and transfer them to the MAC grid (vflip_p2g,
   owner-computes method);
2. keep the transferred grid, add gravity to the v faces;
3. particle level set of the liquid (vflip_liquid_phi, radius RADIUS * h) and
   the free-surface assembly: air cells are Dirichlet p = 0
   (vflip_assemble_liquid), MGPCG to 1e-6, pressure gradient (vflip_grad_apply);
4. particle velocities from the grid change, blended with PIC (vflip_g2p);
5. RK3 particle advection with solid push-out (vflip_advect_particles).

The reference has no liquid solver (its FLIP particles carry smoke); this loop is
the standard FLIP step around the reference's own transfer functions.

    python -m vflip_amd.dambreak [--n 128] [--steps 50] [--ppc 8]
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import api, scenes

# Particle level-set radius in cells: with 8 jittered particles per cell the
# nearest particle is well within 0.9 h of a liquid cell's centre.
RADIUS = 0.9


class DamBreak:
    def __init__(self, n=128, ppc=8, seed=0, flip=0.95, cfl=1.0, gravity=-9.8, max_dt=0.01, tol=1e-6, maxit=100):
        self.n, self.h = n, 1.0 / n
        self.flip, self.cfl, self.gravity, self.max_dt = flip, cfl, gravity, max_dt
        dev = torch.device("cuda")
        h = self.h
        self.solid_phi = torch.from_numpy(scenes.box_phi(n, n, n, h, open_top=False).astype(np.float32)).to(dev)
        self.uw, self.vw, self.ww = api.face_weights(self.solid_phi)
        pos = scenes.block_particles(n, ppc, seed)
        inside = np.all((pos >= 0.5 * h) & (pos < (n - 1.5) * h), axis=1)  # start outside the walls
        self.pos = torch.from_numpy(np.ascontiguousarray(pos[inside])).to(dev)
        self.vel = torch.zeros_like(self.pos)
        self.P = api.Particles(n, n, n, h)
        shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]
        self.grids = [(torch.zeros(s, device=dev), torch.zeros(s, device=dev)) for s in shapes]
        self.old = [torch.zeros(s, device=dev) for s in shapes]
        self.solid_vel = [torch.zeros(s, device=dev) for s in shapes]
        self.lphi = torch.zeros((n, n, n), device=dev)
        self.rhs = torch.zeros((n, n, n), device=dev)
        self.p = torch.zeros((n, n, n), device=dev)
        self.solver = api.Solver(n, n, n, precision=api.FP32, tolerance=tol, max_iterations=maxit)
        self.t = 0.0

    @property
    def count(self):
        return self.pos.shape[0]

    def step(self):
        h = self.h
        vmax = float(self.vel.abs().max()) if self.count else 0.0
        dt = min(self.max_dt, self.cfl * h / vmax) if vmax > 0 else self.max_dt
        _, start, order, _ = self.P.hash(self.pos, count=False)
        (u, _), (v, _), (w, _) = self.P.p2g(self.pos, self.vel, start, order, method=0, out=self.grids)
        for o, g in zip(self.old, (u, v, w)):
            o.copy_(g)
        v.add_(self.gravity * dt)
        api.liquid_phi(h, RADIUS * h, self.pos, start, order, self.lphi.shape, out=self.lphi)
        us, vs, ws = self.solid_vel
        self.solver.assemble_liquid(h, 1.0, u, v, w, self.uw, self.vw, self.ww, us, vs, ws, self.lphi, self.rhs)
        st = self.solver.solve(self.rhs, self.p)
        api.grad_apply(api.COARSE, h, 1.0, self.p, self.uw, self.vw, self.ww, us, vs, ws, u, v, w)
        api.g2p(h, self.pos, self.vel, (u, v, w), self.old, flip=self.flip)
        api.advect_particles(h, dt, self.pos, u, v, w, self.solid_phi)
        self.t += dt
        st["dt"] = dt
        return st


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    sim = DamBreak(n=a.n, ppc=a.ppc, seed=a.seed)
    torch.cuda.synchronize()
    times, its, dts, conv = [], [], [], []
    for s in range(a.steps):
        t0 = time.perf_counter()
        st = sim.step()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        its.append(st["iterations"])
        dts.append(st["dt"])
        conv.append(bool(st["converged"]))
        if s % 10 == 0:
            print(f"step {s}: {times[-1]:.2f} ms, dt {st['dt']:.4g}, its {st['iterations']}", file=sys.stderr,
                  flush=True)
    pos = sim.pos.cpu().numpy()
    out = {"workload": f"FLIP dam break {a.n}^3, {a.ppc} particles/cell in [0,.4]x[0,.6]x[0,1], {a.steps} steps",
           "particles": sim.count, "ms_per_step_mean": float(np.mean(times[1:])), "ms_first_step": times[0],
           "iterations": its, "all_converged": all(conv), "dt_last": dts[-1], "sim_time": sim.t,
           "liquid_centroid": pos.mean(0).tolist(), "liquid_max_y": float(pos[:, 1].max()),
           "finite": bool(np.isfinite(pos).all())}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
