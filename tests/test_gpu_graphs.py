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
