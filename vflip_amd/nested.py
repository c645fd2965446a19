"""Nested-subdomain smoke: the coarse plume plus fine subdomains at 2x refinement.

Each step advances the coarse plume (PlumeSim: advection, forcing, MGPCG
projection), then every fine subdomain the way CoarseFineSmokeSolver couples
its fine grid (S/coarse_fine_solver.cpp):

1. the fine velocity takes the coarse velocity change of the step, sampled at
   the fine faces (CoarseToFineOperator, 146-170), on top of its own detail;
2. semi-Lagrangian RK2 advection on the fine grid;
3. the fine pressure projection with the subdomain's assembly and gradient
   (SubDomainSolver::pcg_projection, S/sub_domain.cpp:556-820: dt-scaled terms,
   ring rows Dirichlet) -- all subdomains' solves either batched (vflip_batch:
   one launch per V-cycle/PCG step, GPU-side stop tests, hipGraph replays,
   autotuned sweep tiles) or one solver call after another;
4. the fine faces averaged back into the coarse faces inside the box
   (FineToCoarseOperator, 73-118).

Subdomain shapes: fine cells per axis drawn from {32, 48, 64, 96} per subdomain
(seeded), boxes placed along the plume column and kept two coarse cells away
from the walls. The solids are the coarse scene's spheres evaluated at the fine
nodes. Fine solves reach the reference's subdomain tolerance, an FP64 true
residual of 1e-9 relative (S/sub_domain.cpp:739): FP64 defect correction rounds
around FP32 MGPCG solves of the corrections (vflip_batch_solve_refined, or
vflip_solver_solve_refined per subdomain), each correction to `inner_tol`.
"""
import gc
import time

import numpy as np
import torch

from . import api, scenes
from .plume import PlumeSim

SIZES = (32, 48, 64, 96)


def layout(n, count, seed, ratio=2):
    """[(c0, m, nf)]: coarse box origin and extent (cells, x/y/z) and fine cells per axis."""
    subs = []
    for q in range(count):
        u = scenes.uniform01(seed, 500 + q, np.arange(6))
        nf = [SIZES[int(u[a] * 4) % 4] for a in range(3)]
        m = [f // ratio for f in nf]
        centre = [n * (0.35 + 0.3 * u[3]), n * (0.12 + 0.76 * (q + 0.5) / count), n * (0.35 + 0.3 * u[5])]
        c0 = [int(min(max(centre[a] - m[a] / 2, 2), n - 2 - m[a])) for a in range(3)]
        subs.append((c0, m, nf))
    return subs


class Subdomain:
    def __init__(self, c0, m, nf, ratio, hc, spheres, dt, inner_tol=1e-5):
        self.c0, self.m, self.nf, self.ratio = c0, m, nf, ratio
        self.hc, self.hf, self.dt = hc, hc / ratio, dt
        nx, ny, nz = nf
        dev = torch.device("cuda")
        f32 = dict(device=dev, dtype=torch.float32)
        self.u = torch.zeros((nz, ny, nx + 1), **f32)
        self.v = torch.zeros((nz, ny + 1, nx), **f32)
        self.w = torch.zeros((nz + 1, ny, nx), **f32)
        self.u2, self.v2, self.w2 = torch.zeros_like(self.u), torch.zeros_like(self.v), torch.zeros_like(self.w)
        self.tmp = [torch.zeros_like(self.u), torch.zeros_like(self.v), torch.zeros_like(self.w)]
        self.us, self.vs, self.ws = torch.zeros_like(self.u), torch.zeros_like(self.v), torch.zeros_like(self.w)
        # solid SDF at the fine nodes: fine node i sits at coarse-frame world ((c0*ratio + i) - 1) * hf
        phi = np.full((nz + 1, ny + 1, nx + 1), 1e3, np.float64)
        gx = ((c0[0] * ratio + np.arange(nx + 1)) - 1.0) * self.hf
        gy = ((c0[1] * ratio + np.arange(ny + 1)) - 1.0) * self.hf
        gz = ((c0[2] * ratio + np.arange(nz + 1)) - 1.0) * self.hf
        for c, r in spheres:
            cc = np.asarray(c, np.float64) - 0.5 * hc  # config -> reference coordinates
            d = np.sqrt((gx[None, None, :] - cc[0]) ** 2 + (gy[None, :, None] - cc[1]) ** 2 +
                        (gz[:, None, None] - cc[2]) ** 2) - r
            phi = np.minimum(phi, d)
        self.phi = torch.from_numpy(phi.astype(np.float32)).to(dev)
        self.uw, self.vw, self.ww = api.face_weights(self.phi)
        self.rhs = torch.zeros((nz, ny, nx), **f32)
        self.p = torch.zeros_like(self.rhs)
        self.p64 = torch.zeros((nz, ny, nx), device=dev, dtype=torch.float64)
        self.solver = api.Solver(nx, ny, nz, precision=api.FP32, tolerance=inner_tol, max_iterations=1000)

    def coarse_box(self, a, comp):
        """The coarse faces of component comp (0/1/2) inside and on this box."""
        (x0, y0, z0), (mx, my, mz) = self.c0, self.m
        ex = [mx, my, mz]
        ex[comp] += 1
        return a[z0:z0 + ex[2], y0:y0 + ex[1], x0:x0 + ex[0]]

    def add_coarse_change(self, coarse_new, coarse_old):
        """fine += coarse (new - old) sampled at the fine faces."""
        for comp, (f, cn, co) in enumerate(zip((self.u, self.v, self.w), coarse_new, coarse_old)):
            d = (self.coarse_box(cn, comp) - self.coarse_box(co, comp)).contiguous()
            foffs = tuple(1.0 if a == comp else 0.5 for a in range(3))
            coffs = tuple(0.5 if a == comp else 0.0 for a in range(3))
            api.coarse_to_fine(self.tmp[comp], foffs, self.hf, d, coffs, self.hc)
            f.add_(self.tmp[comp])

    def advect(self):
        h, dt = float(self.hf), float(self.dt)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.u, (0.5, 0.0, 0.0), self.u2, mode=2)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.v, (0.0, 0.5, 0.0), self.v2, mode=2)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.w, (0.0, 0.0, 0.5), self.w2, mode=2)
        self.u, self.u2 = self.u2, self.u
        self.v, self.v2 = self.v2, self.v
        self.w, self.w2 = self.w2, self.w

    def assemble(self):
        self.solver.assemble(api.SUB, float(self.hf), float(self.dt), self.u, self.v, self.w, self.uw, self.vw,
                             self.ww, self.us, self.vs, self.ws, self.rhs)

    def apply_gradient(self):
        # from the FP64 pressure, as the reference's double pressure (S/sub_domain.cpp:745-819)
        api.grad_apply_f64(api.SUB, float(self.hf), float(self.dt), self.p64, self.uw, self.vw, self.ww, self.us,
                           self.vs, self.ws, self.u, self.v, self.w)

    def write_back(self, coarse):
        """Coarse faces strictly inside the box = mean of the fine faces over them."""
        for comp, (f, c) in enumerate(zip((self.u, self.v, self.w), coarse)):
            box = self.coarse_box(c, comp)
            avg = torch.empty_like(box)
            offs = tuple(self.ratio // 2 if a == comp else 0 for a in range(3))
            api.fine_to_coarse(f, avg, self.ratio, offs)
            sl = [slice(None)] * 3
            sl[2 - comp] = slice(1, box.shape[2 - comp] - 1)
            box[tuple(sl)] = avg[tuple(sl)]


class NestedPlume:
    """Coarse plume of n^3 cells plus `count` fine subdomains at `ratio`x.

    graph: replay the batched PCG steps from a captured hipGraph. Off by default: on
    ROCm 7 replays take the same time as direct launches for these solves, and every
    launch the caller then makes on the legacy default stream waits about 0.45 ms on
    the host (measured: 108 post-solve launches, 49 ms of gaps for 0.46 ms of kernels).
    """

    def __init__(self, n=256, count=12, seed=0, ratio=2, dt=0.01, batched=True, autotune=True, graph=False,
                 fine_tol=1e-9, inner_tol=1e-5, max_rounds=8):
        self.n, self.ratio, self.dt = n, ratio, dt
        self.spheres = scenes.random_spheres(seed)
        phi = scenes.solid_phi(n, 1.0 / n, self.spheres)
        self.coarse = PlumeSim(n, phi, dt=dt)
        self.subs = [Subdomain(c0, m, nf, ratio, 1.0 / n, self.spheres, dt, inner_tol)
                     for c0, m, nf in layout(n, count, seed, ratio)]
        self.fine_tol, self.max_rounds = fine_tol, max_rounds
        self.batched, self.graph = batched, graph
        self.batch = api.Batch([s.solver for s in self.subs]) if batched else None
        self.tiles = None
        self.autotune = autotune
        self.timing = []

    def _solve_fine(self):
        if self.batched:
            if self.autotune and self.tiles is None:
                self.tiles = self.batch.autotune()
            rc, stats, _ = self.batch.solve_refined([s.rhs for s in self.subs], [s.p for s in self.subs],
                                                    tol=self.fine_tol, max_rounds=self.max_rounds,
                                                    x64=[s.p64 for s in self.subs], graph=self.graph)
            return stats
        return [s.solver.solve_refined(s.rhs, s.p, tol=self.fine_tol, max_rounds=self.max_rounds, x64=s.p64)
                for s in self.subs]

    def step(self):
        sync = torch.cuda.synchronize
        old = [t.clone() for t in (self.coarse.u, self.coarse.v, self.coarse.w)]
        sync()
        t0 = time.perf_counter()
        cst = self.coarse.step()
        sync()
        t1 = time.perf_counter()
        new = (self.coarse.u, self.coarse.v, self.coarse.w)
        for s in self.subs:
            s.add_coarse_change(new, old)
            s.advect()
            s.assemble()
        sync()
        t2 = time.perf_counter()
        stats = self._solve_fine()
        sync()
        t3 = time.perf_counter()
        for s in self.subs:
            s.apply_gradient()
            s.write_back(new)
        sync()
        t4 = time.perf_counter()
        rec = {"coarse_ms": (t1 - t0) * 1e3, "fine_prep_ms": (t2 - t1) * 1e3, "fine_solve_ms": (t3 - t2) * 1e3,
               "fine_post_ms": (t4 - t3) * 1e3, "step_ms": (t4 - t0) * 1e3, "coarse_iterations": cst["iterations"],
               "fine_iterations": [st["iterations"] for st in stats],
               "fine_converged": all(st["converged"] for st in stats),
               "fine_max_rel_residual": max(st["rel_residual"] for st in stats),
               "fine_true_residuals": [st["true_residual"] for st in stats]}
        self.timing.append(rec)
        return rec


def run(steps=20, n=256, count=12, seed=0, warmup=2, **kw):
    """Runs the nested workload with batched and with sequential fine solves from the same start."""
    out = {}
    for mode in ("batched", "sequential"):
        sim = NestedPlume(n=n, count=count, seed=seed, batched=(mode == "batched"), **kw)
        for _ in range(warmup):
            sim.step()
        # a full pass of Python's garbage collector over the simulation's objects
        # takes about 10 ms; freezing what exists keeps it out of the timed steps
        # (it had landed in the batched runs' post phase)
        gc.collect()
        gc.freeze()
        sim.timing.clear()
        for _ in range(steps):
            sim.step()
        t = sim.timing
        out[mode] = {
            "step_ms": float(np.mean([r["step_ms"] for r in t])),
            "fine_solve_ms": float(np.mean([r["fine_solve_ms"] for r in t])),
            "coarse_ms": float(np.mean([r["coarse_ms"] for r in t])),
            "fine_prep_ms": float(np.mean([r["fine_prep_ms"] for r in t])),
            "fine_post_ms": float(np.mean([r["fine_post_ms"] for r in t])),
            "fine_iterations_last": t[-1]["fine_iterations"],
            "all_converged": all(r["fine_converged"] for r in t),
            "max_rel_residual": max(r["fine_max_rel_residual"] for r in t),
            "tiles": sim.tiles,
            "p_fine_last": [s.p.detach().cpu().numpy() for s in sim.subs],
        }
        out["subdomains"] = [{"box_origin": c0, "coarse_cells": m, "fine_cells": nf} for c0, m, nf in
                             layout(n, count, seed)]
        del sim
        gc.unfreeze()
        torch.cuda.empty_cache()
    return out
