"""The PCG iteration replayed from captured graphs equals direct launches bit for bit:
repeated solves (graph replays with other work between them), a second right-hand side,
and a re-assembly that changes the operator and the FP16 level scales (recapture)."""
import os

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def _solver(api, n, precision, graphs):
    old = os.environ.get("VFLIP_GRAPHS")
    os.environ["VFLIP_GRAPHS"] = "1" if graphs else "0"
    try:
        return api.Solver(n, n, n, precision=precision, tolerance=1e-6, max_iterations=100)
    finally:
        if old is None:
            del os.environ["VFLIP_GRAPHS"]
        else:
            os.environ["VFLIP_GRAPHS"] = old


@pytest.mark.parametrize("n,prec", [(64, "fp32"), (128, "fp32"), (128, "mixed")])
def test_graph_replay_bitexact(cuda, n, prec):
    print(f"graphs n={n} {prec}", flush=True)
    import torch
    from vflip_amd import api
    precision = api.FP32 if prec == "fp32" else api.MIXED
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    res = {}
    for graphs in (False, True):
        S = _solver(api, n, precision, graphs)
        out = []
        for seed in (1, 2):  # two operators (different sphere): re-assembly, new hierarchy values
            pb = scenes.box_problem(n, n, n, seed)
            g = {k: t(v) for k, v in pb.items() if isinstance(v, np.ndarray)}
            uw, vw, ww = api.face_weights(g["phi"])
            rhs = torch.empty((n, n, n), device="cuda")
            S.assemble(api.COARSE, float(pb["h"]), 1.0, g["u"], g["v"], g["w"], uw, vw, ww, g["us"], g["vs"],
                       g["ws"], rhs)
            for scale in (1.0, -0.5, 2.0):  # replays with eager work between them
                p = torch.zeros((n, n, n), device="cuda")
                st = S.solve(rhs * scale, p)
                torch.cuda.synchronize()
                out.append((p.cpu().numpy(), st["iterations"], st["converged"]))
        res[graphs] = out
        del S
    for (pa, ia, ca), (pb_, ib, cb) in zip(res[False], res[True]):
        assert ia == ib and ca == cb
        assert np.array_equal(pa.view(np.uint32), pb_.view(np.uint32))


def test_half_levels_fused_sweep_bitexact(cuda):
    """FP16-stored levels smoothed by the fused z-marching sweep give the same mixed-precision
    V-cycle and solve, bit for bit, as the per-colour half sweeps. VFLIP_FUSED_MIN_N = 32^3
    fuses every level down to the FP16 32^3 one; 64^3 keeps the FP32 levels (128^3, 64^3)
    fused in both runs, so only the FP16 level changes form (the level-0 sweep also fuses the
    z.r dot, whose summation order would otherwise differ)."""
    print("fp16 fused sweep", flush=True)
    import torch
    from vflip_amd import api
    n = 128
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pb = scenes.box_problem(n, n, n, 4)
    g = {k: t(v) for k, v in pb.items() if isinstance(v, np.ndarray)}
    uw, vw, ww = api.face_weights(g["phi"])
    out = {}
    for fused in (False, True):
        old = os.environ.get("VFLIP_FUSED_MIN_N")
        os.environ["VFLIP_FUSED_MIN_N"] = str(32 ** 3 if fused else 64 ** 3)
        try:
            S = api.Solver(n, n, n, precision=api.MIXED, tolerance=1e-6, max_iterations=100)
        finally:
            if old is None:
                os.environ.pop("VFLIP_FUSED_MIN_N", None)
            else:
                os.environ["VFLIP_FUSED_MIN_N"] = old
        rhs = torch.empty((n, n, n), device="cuda")
        S.assemble(api.COARSE, float(pb["h"]), 1.0, g["u"], g["v"], g["w"], uw, vw, ww, g["us"], g["vs"], g["ws"], rhs)
        z = torch.empty_like(rhs)
        S.vcycle(rhs, z)
        p = torch.zeros_like(rhs)
        st = S.solve(rhs, p)
        torch.cuda.synchronize()
        out[fused] = (z.cpu().numpy(), p.cpu().numpy(), st["iterations"])
        del S
    assert np.array_equal(out[False][0].view(np.uint32), out[True][0].view(np.uint32))
    assert out[False][2] == out[True][2]
    assert np.array_equal(out[False][1].view(np.uint32), out[True][1].view(np.uint32))
