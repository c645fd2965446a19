"""Nested-subdomain workload measurement: batched vs sequential fine solves.

    python -m vflip_amd.bench_nested [--n 256] [--count 12] [--steps 20] [--warmup 2] [--seed 0] [--graph]

Runs the coupled coarse + fine simulation (vflip_amd.nested) twice from the same
start, once with the fine pressure solves batched (vflip_batch: autotuned tiles,
GPU-side stop tests; --graph replays the PCG steps from captured hipGraphs) and
once solver by solver, and prints one JSON line with the
mean per-step times of each phase, the iterations and the difference between
the two runs' final fine pressures.
"""
import argparse
import json
import sys

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--count", type=int, default=12)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay the batched PCG steps from captured hipGraphs")
    a = ap.parse_args(argv)
    from .nested import run
    res = run(steps=a.steps, n=a.n, count=a.count, seed=a.seed, warmup=a.warmup, autotune=not a.no_autotune,
              graph=a.graph)
    pb, ps = res["batched"].pop("p_fine_last"), res["sequential"].pop("p_fine_last")
    rel = max(float(np.linalg.norm(x.astype(np.float64) - y) / max(np.linalg.norm(y.astype(np.float64)), 1e-30))
              for x, y in zip(pb, ps))
    fine_cells = sum(int(np.prod(s["fine_cells"])) for s in res["subdomains"])
    out = {"workload": f"nested smoke: {a.n}^3 coarse + {a.count} fine subdomains at 2x, {a.steps} steps",
           "batched_graph_replay": a.graph,
           "fine_cells": fine_cells, "subdomains": res["subdomains"],
           "batched": res["batched"], "sequential": res["sequential"],
           "fine_solve_speedup": res["sequential"]["fine_solve_ms"] / res["batched"]["fine_solve_ms"],
           "step_speedup": res["sequential"]["step_ms"] / res["batched"]["step_ms"],
           "fine_pressure_rel_l2_batched_vs_sequential": rel}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
