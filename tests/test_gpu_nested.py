"""The nested-subdomain workload (vflip_amd.nested) on a small coarse grid: the
coupled simulation with batched fine solves equals the one with sequential
fine solves bit for bit when the batch keeps the solvers' tile shapes (every
field, every step), and stays within FP32 round-off of it with autotuned tiles.
Every fine solve reaches an FP64 true residual of 1e-9 (FP64 defect correction
around the FP32 MGPCG)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(batched, autotune, steps=3):
    import torch
    from vflip_amd.nested import NestedPlume
    sim = NestedPlume(n=64, count=5, seed=3, batched=batched, autotune=autotune)
    u, v = sim.coarse.u.clone(), sim.coarse.v.clone()
    u.uniform_(-0.5, 0.5, generator=torch.Generator(device="cuda").manual_seed(1))
    v.uniform_(-0.5, 0.5, generator=torch.Generator(device="cuda").manual_seed(2))
    sim.coarse.load_state({"u": u, "v": v})
    recs = [sim.step() for _ in range(steps)]
    fields = [sim.coarse.u, sim.coarse.v, sim.coarse.w, sim.coarse.p] + [t for s in sim.subs for t in (s.u, s.p)]
    return recs, [f.detach().cpu().numpy() for f in fields]


def test_nested_batched_equals_sequential(cuda):
    print("nested: batched (default tiles) vs sequential", flush=True)
    rb, fb = _run(True, False)
    rs, fs = _run(False, False)
    for a, b in zip(rb, rs):
        assert a["fine_iterations"] == b["fine_iterations"] and a["fine_converged"] and b["fine_converged"]
        assert max(a["fine_true_residuals"]) <= 1e-9  # the reference's FP64 subdomain tolerance
    for a, b in zip(fb, fs):
        assert np.array_equal(a, b)
    print("fine iterations per step:", [r["fine_iterations"] for r in rb], flush=True)


def test_nested_autotuned_close(cuda):
    print("nested: batched (autotuned) vs sequential", flush=True)
    rb, fb = _run(True, True)
    rs, fs = _run(False, False)
    assert all(r["fine_converged"] for r in rb)
    for a, b in zip(fb, fs):
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        assert np.linalg.norm(a64 - b64) <= 1e-4 * max(np.linalg.norm(b64), 1e-30)
