"""Training data, training and evaluation of the learned warm start.

Pairs come from the plume workload's own solver at 64^3 (h = 1/64, dt = 0.01):
for each seed, 1-4 seeded random spheres and a seeded random spherical source
(centre in [0.3, 0.7] x [0.05, 0.25] x [0.3, 0.7], radius in [0.05, 0.12],
config coordinates), a seeded number of plume steps (4..15), then one more
advection + forcing and the projection's right-hand side solved in mixed
precision to a 1e-6 FP64 true residual. The network input is (rhs, solid
fraction, liquid fraction); the target is the pressure. Normalisation
(rhs / max|rhs|, pressure / max|rhs|) happens inside the library, so the same
network serves other grids through vflip_unet_warm_start's restriction.

Training runs the library's own kernels (forward, backward, Adam); this module
only orders the samples.
"""
import math
import time

import numpy as np
import torch

from . import api, scenes
from .plume import PlumeSim

N_TRAIN = 64


def scene(seed, n=N_TRAIN):
    """(nodal solid phi, source centre (config coords), source radius, plume steps) for a seed."""
    u = scenes.uniform01(seed, 41, np.arange(8))
    spheres = scenes.random_spheres(seed)
    centre = (0.3 + 0.4 * u[0], 0.05 + 0.2 * u[1], 0.3 + 0.4 * u[2])
    radius = 0.05 + 0.07 * u[3]
    steps = 4 + int(u[4] * 12) % 12
    phi = scenes.solid_phi(n, 1.0 / n, spheres)
    return phi, centre, radius, steps


def generate_pairs(seeds, n=N_TRAIN, log=None):
    """Device tensors rhs, solid, liquid, pressure [S, n, n, n] and the solves' stats."""
    S = len(seeds)
    f32 = dict(device="cuda", dtype=torch.float32)
    rhs = torch.empty((S, n, n, n), **f32)
    solid = torch.empty((S, n, n, n), **f32)
    pres = torch.empty((S, n, n, n), **f32)
    ops = torch.empty((S, 4, n, n, n), **f32)  # the sample's operator: diag, +x, +y, +z couplings
    stats = []
    exact = api.Solver(n, n, n, precision=api.MIXED, tolerance=1e-6, max_iterations=100)
    t0 = time.time()
    for q, seed in enumerate(seeds):
        phi, centre, radius, steps = scene(seed, n)
        sim = PlumeSim(n, phi, dt=0.01, source_center_cfg=centre, source_r=radius, seed=seed)
        for _ in range(steps):
            sim.step()
        sim.advect()
        sim.force()
        st = exact.project(api.COARSE, float(sim.h), 1.0, sim.u, sim.v, sim.w, sim.uw, sim.vw, sim.ww, sim.us, sim.vs,
                           sim.ws, rhs[q], pres[q])
        if st["status"] < 0 or not st["converged"]:
            raise RuntimeError(f"seed {seed}: solve did not converge {st}")
        d, ax, ay, az = exact.get_level(0)
        solid[q] = (d == 0).float()
        for c, a in enumerate((d, ax, ay, az)):
            ops[q, c] = a
        stats.append(st)
        if log and (q + 1) % 200 == 0:
            log(f"generated {q + 1}/{S} pairs in {time.time() - t0:.1f} s")
    return dict(rhs=rhs, solid=solid, liquid=1.0 - solid, p=pres, ops=ops, stats=stats)


def train(net, data, epochs=8, batch=8, lr=1e-3, lr_min=2e-5, seed=0, loss="energy", log=None):
    """Adam over shuffled mini-batches; cosine learning-rate decay. Returns per-epoch mean losses.

    loss "energy": relative A-norm error of the prediction under each sample's
    operator (the quantity that sets a warm-started PCG's iteration count);
    "mse": mean squared error of the normalised pressure."""
    S = data["rhs"].shape[0]
    nbatch = S // batch
    total = epochs * nbatch
    g = torch.Generator().manual_seed(seed)
    step, hist = 0, []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(S, generator=g)
        losses = []
        for b in range(nbatch):
            idx = perm[b * batch:(b + 1) * batch]
            net.zero_grad()
            for q, i in enumerate(idx.tolist()):
                if loss == "energy":
                    o = data["ops"][i]
                    val = net.train_step_energy(data["rhs"][i], data["solid"][i], data["liquid"][i], data["p"][i],
                                                o[0], o[1], o[2], o[3])
                else:
                    val = net.train_step(data["rhs"][i], data["solid"][i], data["liquid"][i], 0.0, data["p"][i])
                losses.append(val)
            step += 1
            a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * (step - 1) / max(total - 1, 1)))
            net.adam(a, step, 1.0 / batch)
        hist.append(sum(losses) / len(losses))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return hist


def evaluate(net, seeds, n=N_TRAIN, factor=1, precision=api.MIXED, tol=1e-6):
    """What the network gains: for each seed's plume state, the same projection
    solved from a zero guess and from the network's guess (both to tol).
    Returns one dict per seed with iterations, solve ms, the guess's relative
    L2 error against the solved pressure and both final true residuals."""
    rows = []
    for seed in seeds:
        phi, centre, radius, steps = scene(seed, n)
        sim = PlumeSim(n, phi, dt=0.01, precision=precision, tol=tol, maxit=100, source_center_cfg=centre,
                       source_r=radius, seed=seed)
        for _ in range(steps):
            sim.step()
        sim.advect()
        sim.force()
        saved = {k: t.clone() for k, t in sim.state().items()}
        st0 = sim.project()
        p_ref = sim.p.clone()
        guess = net.warm_start(sim.rhs, sim.solver.get_level(0)[0], factor)
        sim.load_state(saved)
        sim.use_network(net, factor)
        st1 = sim.project()
        torch.cuda.synchronize()
        rows.append(dict(seed=seed, its_plain=st0["iterations"], its_learned=st1["iterations"],
                         ms_plain=st0["ms_solve"], ms_learned=st1["ms_solve"],
                         guess_rel_err=relative_error(guess, p_ref), true_res_plain=st0["true_residual"],
                         true_res_learned=st1["true_residual"]))
    return rows


def relative_error(pred, p):
    return float((pred - p).norm() / p.norm())


# ------------------------------------------------------------------ history correction
def apply_op(d, ax, ay, az, x):
    """y = A x for a 7-point operator (diag + couplings to +x/+y/+z, as vflip_solver_set_matrix),
    in the dtype of x; arrays shaped (nk, nj, ni)."""
    d, ax, ay, az = (a.to(x.dtype) for a in (d, ax, ay, az))
    y = d * x
    y[:, :, :-1] += ax[:, :, :-1] * x[:, :, 1:]
    y[:, :, 1:] += ax[:, :, :-1] * x[:, :, :-1]
    y[:, :-1, :] += ay[:, :-1, :] * x[:, 1:, :]
    y[:, 1:, :] += ay[:, :-1, :] * x[:, :-1, :]
    y[:-1, :, :] += az[:-1, :, :] * x[1:, :, :]
    y[1:, :, :] += az[:-1, :, :] * x[:-1, :, :]
    return y


def _pool(a, f):
    return torch.nn.functional.avg_pool3d(a[None, None], f)[0, 0]


def generate_correction_pairs(seeds, n=256, factor=4, steps=12, first=3, log=None):
    """Training pairs of the history correction: for each seed's plume (scene(seed) at n^3,
    mixed solves warm-started from the pressure history), every step from `first` on gives
    the divergence the history guess x0 leaves, r0 = b - A x0, and the error e = p - x0 of
    that guess against the converged pressure p; both averaged by `factor` onto the network
    grid, with the solid fraction and the hierarchy's Galerkin operator on that grid (the
    norm of the energy loss). Returns device tensors [S, ...] like generate_pairs."""
    level = int(round(math.log2(factor)))
    m = n // factor
    rows = {k: [] for k in ("rhs", "solid", "p", "ops")}
    stats = []
    t0 = time.time()
    for seed in seeds:
        phi, centre, radius, _ = scene(seed, n)
        sim = PlumeSim(n, phi, dt=0.01, precision=api.MIXED, tol=1e-6, maxit=100, source_center_cfg=centre,
                       source_r=radius, seed=seed)
        sim.use_history(True)
        for s in range(steps):
            sim.advect()
            sim.force()
            if s >= first:
                sim.solver.assemble(api.COARSE, float(sim.h), 1.0, sim.u, sim.v, sim.w, sim.uw, sim.vw, sim.ww,
                                    sim.us, sim.vs, sim.ws, sim.rhs)
                g, used = sim.solver.history_guess(sim.rhs)
            st = sim.project()
            if s < first:
                continue
            if st["status"] < 0 or not st["converged"]:
                raise RuntimeError(f"seed {seed} step {s}: solve did not converge {st}")
            d, ax, ay, az = sim.solver.get_level(0)
            r0 = sim.rhs.double() - apply_op(d, ax, ay, az, g)
            r0[d == 0] = 0
            e = sim.p.double() - g
            e[d == 0] = 0
            rows["rhs"].append(_pool(r0.float(), factor))
            rows["solid"].append(_pool((d == 0).float(), factor))
            rows["p"].append(_pool(e.float(), factor))
            rows["ops"].append(torch.stack(sim.solver.get_level(level)))
            stats.append(dict(seed=seed, step=s, iterations=st["iterations"], history=used))
        del sim
        if log and len(stats) and (seeds.index(seed) + 1) % 20 == 0:
            log(f"correction pairs: {len(stats)} from {seeds.index(seed) + 1} seeds in {time.time() - t0:.1f} s")
    out = {k: torch.stack(v) for k, v in rows.items()}
    assert out["rhs"].shape[1:] == (m, m, m)
    out["liquid"] = 1.0 - out["solid"]
    out["stats"] = stats
    return out


def evaluate_history(net, seeds, n=256, factor=4, steps=12, first=3):
    """What the network gains over the pressure-history guess alone: each seed's plume run
    `steps` steps twice (history only; history + network correction), both mixed to 1e-6;
    per-seed mean iterations and solve ms over the steps from `first` on."""
    rows = []
    for seed in seeds:
        phi, centre, radius, _ = scene(seed, n)
        rec = {"seed": seed}
        for name, use in (("history", False), ("learned", True)):
            sim = PlumeSim(n, phi, dt=0.01, precision=api.MIXED, tol=1e-6, maxit=100, source_center_cfg=centre,
                           source_r=radius, seed=seed)
            if use:
                sim.use_network(net, factor, history=True)
            else:
                sim.use_history(True)
            its, ms, tres = [], [], []
            for s in range(steps):
                st = sim.step()
                if s >= first:
                    its.append(st["iterations"])
                    ms.append(st["ms_solve"])
                    tres.append(st["true_residual"])
            torch.cuda.synchronize()
            rec[f"its_{name}"] = sum(its) / len(its)
            rec[f"ms_{name}"] = sum(ms) / len(ms)
            rec[f"true_res_{name}"] = max(tres)
            del sim
        rows.append(rec)
    return rows
