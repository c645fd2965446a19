#!/usr/bin/env python3
"""Benchmark of the vFLIP per-step hot path on one MI355X.

Workload: smoke plume on a 256^3 MAC grid (h = 1/256, dt = 0.01), 1-4 seeded
sphere obstacles (radius U[0.05,0.15]), sphere source r = 0.08 at
(0.5, 0.1, 0.5) with rho = T = 1 + U[0, 0.05] jitter (seed 0), buoyancy
alpha = 0.1, beta = 0.3, RK2 semi-Lagrangian advection of u, v, w, rho, T and
the MGPCG pressure projection to 1e-6 relative residual (infinity norm,
recursive residual, as the reference). One step = advection + forcing +
projection; inputs are resident in HBM. The headline `value` is ms/step of the
learned warm start + mixed-precision projection path (the solver's pressure-history
guess corrected by the U-Net, which runs on the grid coarsened to 64^3, seeds every
MGPCG solve, which still runs to the tolerance); the history guess without the
network (the non-learned control), mixed precision from a zero guess, the mixed
form the plume configuration words and the plain FP32 MGPCG path are timed beside
it. The mixed run carries the HIP events of the roofline timing below.

The `roofline` object describes the dominant kernel (the fused level-0
red-black smoothing sweep of the mixed-precision solver, which reads the FP16
copy of the level-0 coefficients): algorithmic bytes per launch (20 B/cell)
over the average launch time measured with HIP events on the solver's stream
inside the timed steps of the mixed run (the same kernel on the same 256^3
level as in the headline run, whose timing the events would otherwise add to).
`cpu_baseline` times the FP64 CPU oracle (a restatement of the reference's TBB
path) on one step of the same workload on the host cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
  python bench.py --rows   (each scope row's GPU call beside its CPU oracle path)
  python bench.py --trajectory 20 --n 128   (p/u error against the FP64 oracle after 20 steps)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/step & iters to 1e-6 rel. residual, 256^3: learned+mixed vs plain MGPCG"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scene(n, seed=0):
    from vflip_amd import scenes
    h = 1.0 / n
    spheres = scenes.random_spheres(seed)
    phi = scenes.solid_phi(n, h, spheres)
    return phi, spheres


WEIGHTS = os.path.join(ROOT, "vflip_amd", "data", "unet_correction256.npz")
WEIGHTS_GRID = 256  # the solve grid the correction was trained for (network on 64^3)


def load_network():
    """The trained correction U-Net of the history warm start (weights committed with the
    library), or None."""
    from vflip_amd.unet import UNet
    if not os.path.exists(WEIGHTS):
        return None
    net = UNet(seed=0)
    net.load(WEIGHTS)
    net.trained_gain = net.gain
    return net


def configure_network(net, n):
    """Pressure unit for a solve grid of n^3: the correction scales with h^2 (the operator
    carries 1/h^2), so the gain trained at WEIGHTS_GRID is rescaled by (WEIGHTS_GRID / n)^2."""
    net.gain = net.trained_gain * (WEIGHTS_GRID / n) ** 2
    return net


def run_gpu(n, steps, warmup, precision, phi, init_state=None, profile=False, net=None, solver_opts=None,
            history=False):
    import torch
    from vflip_amd import api
    from vflip_amd.plume import PlumeSim
    sim = PlumeSim(n, phi, precision=precision)
    if solver_opts:
        sim.solver.set_opts(**solver_opts)
    if net is not None:
        configure_network(net, n)
        sim.use_network(net, max(1, n // 64), history=history)
    elif history:
        sim.use_history(True)
    if init_state is not None:
        sim.load_state(init_state)
    for _ in range(warmup):
        sim.step()
    torch.cuda.synchronize()
    if profile:
        sim.solver.profile(True)
    its, proj_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = sim.step()
        its.append(st["iterations"])
        proj_ms.append(st["ms_solve"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = sim.solver.profile_read() if profile else None
    if profile:
        sim.solver.profile(False)
    return sim, (t1 - t0) * 1e3 / steps, its, proj_ms, prof, st


def traffic_from_profiles(kernel):
    """HBM bytes per launch of a row-wave kernel at 256^3 from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", "row_kernels_pmc_256.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["kernels"][kernel]["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sim_state, phi, n, dt=0.01, reps=5):
    """The same plume step on the FP64 CPU oracle (all host cores): one warm-up
    step, then the median of `reps` timed steps from the same state."""
    import numpy as np
    runs = [_cpu_step(sim_state, phi, n, dt) for _ in range(1 + reps)]
    timed = runs[1:]
    out = dict(runs[0])
    out["ms_per_step"] = float(np.median([r["ms_per_step"] for r in timed]))
    out["ms_solve"] = float(np.median([r["ms_solve"] for r in timed]))
    out["reps"] = reps
    return out


def _cpu_step(sim_state, phi, n, dt):
    import numpy as np
    from oracle import oracle as O
    h = np.float32(1.0 / n)
    st = {k: (v if isinstance(v, np.ndarray) else v.detach().cpu().numpy()) for k, v in sim_state.items()}
    uw, vw, ww = O.face_weights(phi)
    z = [np.zeros_like(st[k]) for k in ("u", "v", "w")]
    t0 = time.perf_counter()
    u2 = O.advect_rk2(h, np.float32(dt), st["u"], st["v"], st["w"], st["u"], (0.5, 0, 0))
    v2 = O.advect_rk2(h, np.float32(dt), st["u"], st["v"], st["w"], st["v"], (0, 0.5, 0))
    w2 = O.advect_rk2(h, np.float32(dt), st["u"], st["v"], st["w"], st["w"], (0, 0, 0.5))
    r2 = O.advect_rk2(h, np.float32(dt), st["u"], st["v"], st["w"], st["rho"], (0, 0, 0))
    t2 = O.advect_rk2(h, np.float32(dt), st["u"], st["v"], st["w"], st["temp"], (0, 0, 0))
    c = [x - 0.5 * float(h) for x in (0.5, 0.1, 0.5)]
    r2, t2, v2 = O.plume_force(float(h), dt, c, 0.08, 0, 0.05, 0.1, 0.3, r2, t2, v2)
    d, ax, ay, az, rhs = O.assemble(h, u2, v2, w2, uw, vw, ww, *z)
    t_asm = time.perf_counter()
    out = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    t_solve = time.perf_counter()
    uo, vo, wo = O.grad_apply(h, out["x"], uw, vw, ww, *z, u2, v2, w2)
    t1 = time.perf_counter()
    return {
        "ms_per_step": (t1 - t0) * 1e3,
        "ms_solve": (t_solve - t_asm) * 1e3,
        "iterations": out["iterations"],
        "p": out["x"], "u": uo, "v": vo, "w": wo, "rho": r2, "temp": t2,
    }


def trajectory(n, steps, phi, spheres, paths=("plain", "mixed", "mixed_history", "learned_mixed")):
    """Error of the GPU plume after `steps` steps from the initial state against
    the FP64 CPU oracle run for the same steps: relative L2 of p, u, v, w and rho
    for each of the plain, mixed and learned + mixed paths, with per-step
    iterations and FP64 true residuals."""
    import numpy as np
    import torch
    from vflip_amd import api
    from vflip_amd.plume import PlumeSim
    rel = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))
    init = {k: t.detach().cpu().numpy().copy() for k, t in PlumeSim(n, phi, precision=api.FP32).state().items()}
    log(f"trajectory: {steps} oracle steps at {n}^3 ...")
    st = dict(init)
    its_cpu = []
    for s in range(steps):
        o = _cpu_step(st, phi, n, 0.01)
        st = {"u": o["u"], "v": o["v"], "w": o["w"], "rho": o["rho"], "temp": o["temp"], "p": o["p"]}
        its_cpu.append(o["iterations"])
        log(f"oracle step {s}: {o['ms_per_step']:.0f} ms, {o['iterations']} its")
    todo = []
    for name in paths:
        if name == "plain":
            todo.append((name, api.FP32, None))
        elif name == "mixed":
            todo.append((name, api.MIXED, None))
        elif name == "mixed_history":
            todo.append((name, api.MIXED, "history"))
        elif name == "learned_mixed":
            net = load_network()
            if net is not None:
                todo.append((name, api.MIXED, configure_network(net, n)))
    out = {"workload": f"smoke plume {n}^3 ({len(spheres)} sphere obstacles), {steps} steps from the initial state",
           "oracle": "FP64 CPU restatement, AMGPCG to 1e-6", "oracle_iterations": its_cpu, "paths": {}}
    for name, prec, nt in todo:
        sim = PlumeSim(n, phi, precision=prec)
        if nt == "history":
            sim.use_history(True)
        elif nt is not None:
            sim.use_network(nt, max(1, n // 64), history=True)
        its, tres = [], []
        for _ in range(steps):
            r = sim.step()
            its.append(r["iterations"])
            tres.append(r["true_residual"])
        torch.cuda.synchronize()
        g = {k: t.detach().cpu().numpy() for k, t in sim.state().items()}
        out["paths"][name] = {"iterations": its, "true_residual_max": max(tres),
                              **{f"{k}_rel_l2": rel(g[k], st[k]) for k in ("p", "u", "v", "w", "rho")}}
        log(f"{name}: {out['paths'][name]}")
        del sim
    return out


def _gpu_ms(fn, reps=10):
    """Median event-timed duration (ms) of fn() on the current stream, after one untimed call."""
    import numpy as np
    import torch
    fn()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def _cpu_ms(fn):
    t0 = time.perf_counter()
    out = fn()
    return (time.perf_counter() - t0) * 1e3, out


def row_baselines():
    """Each scope row's GPU call and the CPU oracle (the reference algorithm restated,
    OpenMP where the reference's loop is a TBB parallel_for) on the same seeded input.
    Returns a list of dicts; GPU times are medians of 10 event-timed calls."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from vflip_amd import api, scenes
    cores = O.lib().or_num_threads()
    rows = []
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def add(row, workload, gpu_ms, cpu_ms, cpu_cores, parity, alg_bytes=None, bytes_note=None, **extra):
        r = {"row": row, "workload": workload, "gpu_ms": round(gpu_ms, 4), "cpu_ms": round(cpu_ms, 1),
             "cpu_cores": cpu_cores, "cpu_kind": "port", "gpu_over_cpu": round(cpu_ms / gpu_ms, 1), "parity": parity}
        if alg_bytes:
            gbs = alg_bytes / (gpu_ms * 1e-3) / 1e9
            r["roofline"] = {"bound": "hbm", "algorithmic_bytes": int(alg_bytes), "bytes_per_unit": bytes_note,
                             "achieved_GBps": round(gbs, 1), "peak_GBps": HBM_PEAK_GBS,
                             "frac": round(gbs / HBM_PEAK_GBS, 4)}
        r.update(extra)
        rows.append(r)
        log(f"row {row}: gpu {gpu_ms:.3f} ms, cpu {cpu_ms:.1f} ms ({cpu_cores} cores)")

    n = 128
    h = np.float32(1.0 / n)
    pb = scenes.box_problem(n, n, n, 3)
    g = {k: t(v) for k, v in pb.items() if isinstance(v, np.ndarray)}
    uw, vw, ww = api.face_weights(g["phi"])
    # a14 advection of one face component
    out = torch.empty_like(g["u"])
    gm0 = _gpu_ms(lambda: api.advect_rk2(float(h), 0.01, g["u"], g["v"], g["w"], g["u"], (0.5, 0, 0), out, mode=0))
    gm2 = _gpu_ms(lambda: api.advect_rk2(float(h), 0.01, g["u"], g["v"], g["w"], g["u"], (0.5, 0, 0), out, mode=2))
    cm, _ = _cpu_ms(lambda: O.advect_rk2(h, np.float32(0.01), pb["u"], pb["v"], pb["w"], pb["u"], (0.5, 0, 0)))
    ncell = n ** 3
    add("a14 advect_field (RK2, trilinear)", f"u component, {n}^3", gm0, cm, cores, "bit-exact (FP64 mode 0)",
        16 * ncell, "16 B/cell: u, v, w read, one component written", gpu_ms_fp32_mode2=round(gm2, 4))
    # a7 assembly, a13 gradient
    S = api.Solver(n, n, n, precision=api.FP32)
    rhs = torch.empty((n, n, n), device="cuda")
    ga = _gpu_ms(lambda: S.assemble(api.COARSE, float(h), 1.0, g["u"], g["v"], g["w"], uw, vw, ww, g["us"], g["vs"],
                                    g["ws"], rhs))
    uwn, vwn, wwn = O.face_weights(pb["phi"])
    ca, asm = _cpu_ms(lambda: O.assemble(h, pb["u"], pb["v"], pb["w"], uwn, vwn, wwn, pb["us"], pb["vs"], pb["ws"]))
    add("a7 pcg_projection assembly", f"{n}^3 box + sphere", ga, ca, cores, "within one FP32 rounding of FP64",
        72 * ncell, "72 B/cell: velocity, weights, solid velocity read; d, ax, ay, az, rhs and FP64 diag/rhs written")
    pres = np.random.default_rng(0).standard_normal((n, n, n))
    gp = t(pres.astype(np.float32))
    uu, vv, wv = g["u"].clone(), g["v"].clone(), g["w"].clone()
    gg = _gpu_ms(lambda: api.grad_apply(api.COARSE, float(h), 1.0, gp, uw, vw, ww, g["us"], g["vs"], g["ws"], uu, vv,
                                        wv))
    cg, _ = _cpu_ms(lambda: O.grad_apply(h, pres, uwn, vwn, wwn, pb["us"], pb["vs"], pb["ws"], pb["u"], pb["v"],
                                         pb["w"]))
    add("a13 pressure gradient", f"{n}^3", gg, cg, cores, "bit-exact (FP64 pressure)",
        52 * ncell, "52 B/cell: p, weights, solid velocity read; u, v, w read and written")
    del S
    # a17 binning and a18 rebuild: the FLIP smoke set, 8 per cell over 1.6 cells, shuffled
    pos = scenes.block_particles(n, 8, 21, hi=(1.0, 1.0, 1.0), jitter=1.6)
    pos = np.ascontiguousarray(pos[np.random.default_rng(21).permutation(len(pos))])
    vel = scenes.particle_velocities(len(pos), 21)
    P = api.Particles(n, n, n, float(h))
    tp, tv = t(pos), t(vel)
    gh = _gpu_ms(lambda: P.hash(tp, count=False))
    ch, (rk, rs, ro) = _cpu_ms(lambda: O.bin_particles(pos, h, (n, n, n)))
    add("a17 hash_FLIP", f"{len(pos) / 1e6:.1f}M particles, {n}^3, random order", gh, ch, 1, "bit-exact",
        72 * len(pos), "72 B/particle (counting sort: position, key, permutation, index)")
    grids = scenes.random_mac(n, 5, scale=1.0)
    phi = scenes.solid_phi(n, 1.0 / n, [((0.5, 0.5, 0.5), 0.15)], open_top=False)
    tg = [t(a) for a in grids]
    tphi = t(phi)
    _, start, order, _ = P.hash(tp)
    bufs = (torch.empty((len(pos) + 8 * n ** 3, 3), device="cuda"), torch.empty((len(pos) + 8 * n ** 3, 3),
                                                                                  device="cuda"))
    gr = _gpu_ms(lambda: P.reorder(tp, tv, start, order, *tg, tphi, 1, out=bufs))
    cr, _ = _cpu_ms(lambda: O.reorder(pos, vel, h, (n, n, n), rs, ro, *grids, phi, 1))
    add("a18 reorder_FLIP", f"{len(pos) / 1e6:.1f}M particles, {n}^3, random order, box + sphere", gr, cr, 1,
        "bit-exact", 28 * len(pos) + 24 * len(pos) + 4 * ncell,
        "28 B per binned particle read, 24 B per output particle written, 4 B per cell")
    del tp, tv, bufs
    # a19 P2G, a20 G2P, a15 particle advection: the dam-break fill
    pos = scenes.block_particles(n, 8, 0)
    vel = scenes.particle_velocities(len(pos), 0)
    tp, tv = t(pos), t(vel)
    _, start, order, _ = P.hash(tp)
    _, rs, ro = O.bin_particles(pos, h, (n, n, n))
    gp0 = _gpu_ms(lambda: P.p2g(tp, tv, start, order, method=0))
    gp1 = _gpu_ms(lambda: P.p2g(tp, tv, start, order, method=1), reps=3)
    specs = ((0, (n, n, n + 1), (0.5, 0, 0)), (1, (n, n + 1, n), (0, 0.5, 0)), (2, (n + 1, n, n), (0, 0, 0.5)))
    cp, _ = _cpu_ms(lambda: [O.p2g(pos, vel, c, h, (n, n, n), rs, ro, fs, of) for c, fs, of in specs])
    add("a19 particle_to_grid", f"{len(pos) / 1e6:.2f}M particles (dam-break fill), {n}^3, 3 components", gp0, cp,
        cores, "FP32 rounding (owner tiles); exact-order method bit-exact", 24 * len(pos) + 24 * ncell,
        "24 B/particle read + 24 B/cell written (3 fields, 3 weights)", gpu_ms_exact_order=round(gp1, 3))
    du = [a * np.float32(0.1) for a in grids]
    tdu = [t(a) for a in du]
    vv = tv.clone()
    gq = _gpu_ms(lambda: api.g2p(float(h), tp, vv, tdu))
    cq, _ = _cpu_ms(lambda: O.g2p(pos, vel, h, (n, n, n), *du))
    add("a20 particle_from_grid", f"{len(pos) / 1e6:.2f}M particles, {n}^3", gq, cq, cores, "bit-exact",
        36 * len(pos) + 12 * ncell, "24 B read + 12 B written per particle, 12 B/cell (FP64 samplers: compute-bound)")
    pp = tp.clone()
    gv = _gpu_ms(lambda: api.advect_particles(float(h), 0.01, pp, *tg, tphi))
    cv, _ = _cpu_ms(lambda: O.advect_particles(pos, h, np.float32(0.01), *grids, phi))
    add("a15 advect_FLIP (RK3 + push-out)", f"{len(pos) / 1e6:.2f}M particles, {n}^3", gv, cv, cores, "bit-exact",
        24 * len(pos) + 16 * ncell, "24 B/particle, 16 B/cell (9 FP64 samples per particle: compute-bound)")
    # C3 particle level set
    lphi = torch.empty((n, n, n), device="cuda")
    gl = _gpu_ms(lambda: api.liquid_phi(float(h), 0.9 * float(h), tp, start, order, (n, n, n), out=lphi))
    cl, _ = _cpu_ms(lambda: O.liquid_phi(pos, h, (n, n, n), np.float32(0.9) * h))
    add("C3 particle level set", f"{len(pos) / 1e6:.2f}M particles, {n}^3", gl, cl, 1, "bit-exact",
        16 * len(pos) + 12 * ncell, "16 B/particle (position, order), 12 B/cell (starts, phi)")
    del P, tp, tv
    # nested coupling: ghosts and PSE viscosity
    coarse, sub = scenes.nested_flip_case(n=n, box=(0.3, 0.7), seed=1, solid_vel=(0.2, -0.1, 0.05))
    dev = lambda d: {k: (t(v) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
    C, Sd = dev(coarse), dev(sub)
    m = sub["u"].shape[0]
    Ps = api.Particles(m, m, m, float(sub["h"]))
    gs = _gpu_ms(lambda: api.seed_ghost(Ps, C, Sd, 3))
    cs, (gpos, gvel) = _cpu_ms(lambda: O.seed_ghost(coarse, sub, 3))
    add("seed_ghost", f"{m}^3 subdomain at 2x, sphere", gs, cs, 1, "bit-exact")
    rng = np.random.default_rng(3)
    real = (np.asarray(sub["bmin"], np.float32) + rng.uniform(0, m * float(sub["h"]), (8 * m ** 3, 3))).astype(
        np.float32)
    rvel = rng.standard_normal(real.shape).astype(np.float32)
    allp, allv = np.concatenate([real, gpos]), np.concatenate([rvel, gvel])
    tap, tav = t(allp), t(allv)
    _, sst, sor, _ = Ps.hash(tap, clamp=True, bmin=sub["bmin"])
    _, rs, ro = O.bin_particles(allp, np.float32(sub["h"]), (m, m, m), clamp=True, bmin=sub["bmin"])
    gd = _gpu_ms(lambda: api.pse_diffuse(C, Sd, 0.01, 0.05, len(real), tap, tav, sst, sor))
    cd, _ = _cpu_ms(lambda: O.pse_diffuse(coarse, sub, np.float32(0.01), np.float32(0.05), len(real), allp, allv, rs,
                                          ro))
    add("diffuse (PSE viscosity)", f"{len(real) / 1e6:.2f}M subdomain particles + {len(gpos) / 1e6:.2f}M ghosts",
        gd, cd, cores, "within 1e-6 (device exp)")
    # smoke tracers: RK3 through the blended coarse/subdomain velocity, push-out, decay
    rng = np.random.default_rng(9)
    ntr = 1 << 20
    trs = np.concatenate([rng.uniform(0.05, 0.95, (ntr, 3)), rng.uniform(0.5, 1.0, (ntr, 1))], 1).astype(np.float32)
    ttr = t(trs)
    gt = _gpu_ms(lambda: api.advect_tracers(C, Sd, 0.01, 0.2, ttr))
    ct, _ = _cpu_ms(lambda: O.advect_tracers(coarse, sub, np.float32(0.01), np.float32(0.2), trs))
    add("advect_tracers (with subdomain)", f"{ntr / 1e6:.2f}M tracers, {n}^3 coarse + {m}^3 subdomain", gt, ct, cores,
        "bit-exact", 32 * ntr, "16 B read + 16 B written per tracer (FP64 nested samples: compute-bound)")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--rows", action="store_true",
                    help="instead of the headline step: each scope row's GPU call beside its CPU oracle path")
    ap.add_argument("--trajectory", type=int, default=0, metavar="K",
                    help="instead of the headline step: p/u error against the FP64 oracle after K steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from vflip_amd import api
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.rows:
        print(json.dumps({"rows": row_baselines()}), flush=True)
        return
    n = args.n
    phi, spheres = make_scene(n)
    log(f"scene n={n} spheres={spheres}")
    if args.trajectory > 0:
        print(json.dumps(trajectory(n, args.trajectory, phi, spheres)), flush=True)
        return

    # plain FP32 MGPCG from a zero guess (the reference algorithm)
    sim_p, ms_plain, its_plain, proj_plain, _, st_p = run_gpu(n, args.steps, args.warmup, api.FP32, phi)
    log(f"plain: {ms_plain:.3f} ms/step its={its_plain}")
    state0 = {k: t.clone() for k, t in sim_p.state().items()}
    sims = {"plain": sim_p}
    # mixed precision from a zero guess; the dominant kernel is timed in this run
    sim_m, ms_mixed, its_mixed, proj_mixed, prof, st = run_gpu(n, args.steps, args.warmup, api.MIXED, phi,
                                                               profile=True)
    log(f"mixed: {ms_mixed:.3f} ms/step its={its_mixed}")
    sims["mixed"] = sim_m
    # the mixed form as the plume configuration words it: FP16 copies of levels >= 2 only,
    # FP64 true-residual refresh every 8 iterations (timed beside the tuned default)
    cfg_opts = {"half_from_level": 2, "refine_every": 8}
    sim_c, ms_cfg, its_cfg, proj_cfg, _, st_c = run_gpu(n, args.steps, args.warmup, api.MIXED, phi,
                                                        solver_opts=cfg_opts)
    log(f"mixed (configuration form): {ms_cfg:.3f} ms/step its={its_cfg}")
    mixed_cfg = {"ms_per_step": round(ms_cfg, 3), "iterations": its_cfg,
                 "projection_ms": round(float(np.mean(proj_cfg)), 3), "true_residual_last": st_c["true_residual"],
                 "solver_opts": cfg_opts}
    del sim_c
    # the non-learned control: mixed warm-started from the solver's pressure history alone
    sim_h, ms_hist, its_hist, proj_hist, _, st_h = run_gpu(n, args.steps, args.warmup, api.MIXED, phi, history=True)
    log(f"mixed + pressure history: {ms_hist:.3f} ms/step its={its_hist}")
    sims["mixed_history"] = sim_h
    history = {"ms_per_step": round(ms_hist, 3), "iterations": its_hist,
               "projection_ms": round(float(np.mean(proj_hist)), 3), "true_residual_last": st_h["true_residual"]}
    # learned + mixed (headline): the history guess corrected by the U-Net from the divergence it
    # leaves, on the grid coarsened to 64^3
    net = load_network()
    learned = None
    st_l = None
    if net is not None:
        sim_l, ms_learned, its_learned, proj_learned, _, st_l = run_gpu(n, args.steps, args.warmup, api.MIXED,
                                                                        phi, net=net, history=True)
        log(f"learned+mixed: {ms_learned:.3f} ms/step its={its_learned}")
        learned = {"ms_per_step": round(ms_learned, 3), "iterations": its_learned,
                   "projection_ms": round(float(np.mean(proj_learned)), 3),
                   "true_residual_last": st_l["true_residual"], "network_grid": n // max(1, n // 64),
                   "factor": max(1, n // 64), "weights": os.path.relpath(WEIGHTS, ROOT),
                   "network_gain_ms_per_step_vs_history": round(ms_hist - ms_learned, 3),
                   "network_gain_iterations_per_step_vs_history": round(float(np.mean(its_hist) -
                                                                              np.mean(its_learned)), 3)}
        sims["learned_mixed"] = sim_l

    roofline = None
    if prof and prof["launches"] > 0:
        avg_ms = prof["total_ms"] / prof["launches"]
        achieved = prof["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profiles("k_rbgs_zrow<4, 0, double, __half>") if n == 256 else None,
                    "kernel": "k_rbgs_zrow<4, 0, double, __half> (fused red-black sweep, level 0, mixed run)",
                    "avg_launch_us": round(avg_ms * 1e3, 2), "launches": prof["launches"],
                    "algorithmic_bytes_per_launch": prof["bytes_per_launch"]}

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        # one step from the plain run's final state on the CPU oracle, and the same step on each GPU path
        log("cpu baseline: FP64 oracle steps (1 warm-up + 5 timed) ...")
        base = cpu_baseline(state0, phi, n)
        from oracle import oracle as O
        cores = O.lib().or_num_threads()
        cpu = {"value": round(base["ms_per_step"], 1), "unit": "ms/step", "cores": cores, "kind": "port",
               "sample": f"one plume step at {n}^3 (advection of u,v,w,rho,T + forcing + assembly + FP64 AMGPCG "
                         f"to 1e-6, {base['iterations']} its + gradient), FP64 oracle with OpenMP; median of "
                         f"{base['reps']} after 1 warm-up",
               "ms_solve": round(base["ms_solve"], 1), "cpu_model": cpu_model(), "nproc": os.cpu_count()}

        def rel(a, b):
            a = a.detach().cpu().numpy().astype(np.float64)
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        oracle_err = {"oracle_iterations": base["iterations"]}
        for name, g in sims.items():
            g.load_state(state0)
            stg = g.step()
            torch.cuda.synchronize()
            oracle_err[name] = {"p_rel_l2": rel(g.p, base["p"]), "u_rel_l2": rel(g.u, base["u"]),
                                "v_rel_l2": rel(g.v, base["v"]), "w_rel_l2": rel(g.w, base["w"]),
                                "gpu_iterations": stg["iterations"], "true_residual": stg["true_residual"]}
        log(f"cpu: {cpu} err: {oracle_err}")
    del sims

    voxels = n ** 3
    head = learned["ms_per_step"] if learned else ms_hist
    dflt = api.default_opts()
    out = {
        "metric": METRIC,
        "value": round(head, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"smoke plume {n}^3 (1-4 seeded sphere obstacles), RK2 semi-Lagrangian advection + "
                               f"MGPCG projection to 1e-6; headline = learned warm start (pressure-history guess "
                               f"corrected by the U-Net) + mixed-precision projection" +
                               ("" if learned else " (weights missing: history guess only)"),
                   "grid": n, "dt": 0.01, "tolerance": 1e-6, "spheres": len(spheres)},
        "iterations": {"learned_mixed": learned["iterations"] if learned else None, "mixed_history": its_hist,
                       "mixed": its_mixed, "plain": its_plain},
        "plain_ms_per_step": round(ms_plain, 3),
        "mixed_ms_per_step": round(ms_mixed, 3),
        "mixed_history_ms_per_step": round(ms_hist, 3),
        "learned_mixed_ms_per_step": learned["ms_per_step"] if learned else None,
        "projection_ms": {"plain": round(float(np.mean(proj_plain)), 3), "mixed": round(float(np.mean(proj_mixed)), 3),
                          "mixed_history": history["projection_ms"],
                          "learned_mixed": learned["projection_ms"] if learned else None},
        "speedup_vs_plain": round(ms_plain / head, 3),
        "speedup_mixed_vs_plain": round(ms_plain / ms_mixed, 3),
        "speedup_mixed_history_vs_plain": round(ms_plain / ms_hist, 3),
        "voxels_per_s": round(voxels / (head * 1e-3), 1),
        "true_residual_last": learned["true_residual_last"] if learned else st_h["true_residual"],
        # last solve of each run: recursive |r|_inf/|b|_inf (the stop test) and the FP64 true residual
        "residuals_last": {"plain": {"recursive": st_p["rel_residual"], "true_fp64": st_p["true_residual"]},
                           "mixed": {"recursive": st["rel_residual"], "true_fp64": st["true_residual"]},
                           "mixed_history": {"recursive": st_h["rel_residual"], "true_fp64": st_h["true_residual"]},
                           "learned_mixed": ({"recursive": st_l["rel_residual"], "true_fp64": st_l["true_residual"]}
                                             if learned else None)},
        "learned_warm_start": learned,
        "history_warm_start": history,
        # the tuned mixed form above (FP16 coefficients at every V-cycle level, FP64 refresh
        # every 4 iterations) beside the form the plume configuration words
        "mixed_solver_opts": {"half_from_level": dflt.half_from_level, "refine_every": dflt.refine_every},
        "mixed_configuration_form": mixed_cfg,
        "fp64_oracle_error": oracle_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
