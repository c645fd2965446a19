"""Batched subdomain pressure solves (vflip_batch) against the same solvers run
one by one (vflip_solver_solve).

With the solvers' own tile shapes a batched solve executes, per subdomain, the
same kernels with the same block decomposition in the same order, so the
solutions, iteration counts and stop decisions are bit-identical to the
sequential solves: eager launches and the captured hipGraph alike, over several
graph replays with eager work and synchronisations between them. Autotuned tile
shapes regroup the z.r partial sums, so those solutions agree to FP32 round-off
(1e-5 relative L2) at the same tolerance. Subdomain shapes are uneven (flat and
fused level-0 smoothers, 2 to 4 levels), one right-hand side is zero and one
subdomain hits its iteration limit.
"""
import numpy as np
import pytest

from vflip_amd import scenes

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (48, 32, 64), (64, 48, 32), (20, 24, 28), (96, 32, 40), (40, 56, 48)]


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _systems(api, shapes, seed0, **opts):
    """One FP32 solver per shape with its subdomain operator assembled; returns solvers and right-hand sides."""
    import torch
    solvers, rhs = [], []
    for q, (ni, nj, nk) in enumerate(shapes):
        pb = scenes.box_problem(ni, nj, nk, seed0 + q)
        S = api.Solver(ni, nj, nk, tolerance=1e-6, max_iterations=100, **opts)
        uw, vw, ww = api.face_weights(_t(pb["phi"]))
        r = torch.zeros((nk, nj, ni), device="cuda")
        S.assemble(api.SUB, float(pb["h"]), 0.01, _t(pb["u"]), _t(pb["v"]), _t(pb["w"]), uw, vw, ww, _t(pb["us"]),
                   _t(pb["vs"]), _t(pb["ws"]), r)
        solvers.append(S)
        rhs.append(r)
    return solvers, rhs


def _sequential(solvers, rhs):
    import torch
    xs, sts = [], []
    for S, r in zip(solvers, rhs):
        x = torch.zeros_like(r)
        sts.append(S.solve(r, x))
        xs.append(x.cpu().numpy())
    return xs, sts


def test_batch_matches_sequential_bitexact(cuda):
    print("batch vs sequential, eager and graph", flush=True)
    import torch
    from vflip_amd import api
    solvers, rhs = _systems(api, SHAPES, 10)
    rhs[3].zero_()  # |b| = 0: converged at once with x = 0
    B = api.Batch(solvers)
    B.set_graph_iterations(4)
    for rep, (graph, scale) in enumerate([(False, 1.0), (True, 1.0), (True, 0.5), (True, 3.0)]):
        rr = [r * scale for r in rhs]
        xs_ref, st_ref = _sequential(solvers, rr)
        xb = [torch.full_like(r, float("nan")) for r in rr]
        rc, st, ms = B.solve(rr, xb, graph=graph)
        torch.cuda.synchronize()
        assert rc == 0, st
        for q in range(len(SHAPES)):
            assert st[q]["iterations"] == st_ref[q]["iterations"], (rep, q)
            assert st[q]["converged"] == 1
            assert np.array_equal(xb[q].cpu().numpy(), xs_ref[q]), (rep, q)
            assert st[q]["rel_residual"] == st_ref[q]["rel_residual"]
        assert st[3]["iterations"] == 0 and not xb[3].any()
        print(f"rep {rep} graph={graph}: iterations {[s['iterations'] for s in st]}, {ms:.2f} ms", flush=True)


def test_batch_iteration_limit(cuda):
    print("batch iteration limit", flush=True)
    import torch
    from vflip_amd import api
    solvers, rhs = _systems(api, SHAPES[:3], 30)
    solvers[1].set_opts(max_iterations=2)
    B = api.Batch(solvers)
    xs_ref, st_ref = _sequential(solvers, rhs)
    xb = [torch.zeros_like(r) for r in rhs]
    rc, st, _ = B.solve(rhs, xb, graph=True)
    assert rc == 1
    assert st[1]["converged"] == 0 and st[1]["iterations"] == 2 == st_ref[1]["iterations"]
    for q in range(3):
        assert np.array_equal(xb[q].cpu().numpy(), xs_ref[q])
    assert st[0]["converged"] == 1 and st[2]["converged"] == 1


def test_batch_autotuned(cuda):
    print("batch autotune", flush=True)
    import torch
    from vflip_amd import api
    solvers, rhs = _systems(api, SHAPES, 50)
    B = api.Batch(solvers)
    shapes = B.autotune()
    print("tuned (rows, z chunk) per level depth:", shapes, flush=True)
    assert shapes[0][0] in (4, 8, 16) and shapes[0][1] in (4, 8, 16, 32, 64)
    xs_ref, st_ref = _sequential(solvers, rhs)
    xb = [torch.zeros_like(r) for r in rhs]
    rc, st, _ = B.solve(rhs, xb, graph=True)
    assert rc == 0
    for q in range(len(SHAPES)):
        ref = xs_ref[q].astype(np.float64)
        got = xb[q].cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
        assert abs(st[q]["iterations"] - st_ref[q]["iterations"]) <= 1
        assert st[q]["rel_residual"] <= 1e-6
