"""GPU advection and plume forcing against the oracle: bit-exact (same FP64
interpolation, same float/double rounding points, no FMA contraction)."""
import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n,mode", [(24, 0), (33, 0), (24, 1), (40, 1)])
def test_advect_bitexact(cuda, n, mode):
    from oracle import oracle as O
    from vflip_amd import api
    print(f"advect n={n} mode={mode}", flush=True)
    h = np.float32(1.0 / n)
    dt = np.float32(0.6 / n)  # CFL ~ 0.6 at |u| ~ 1
    u, v, w = scenes.random_mac(n, 17, scale=1.0)
    rng = np.random.default_rng(n)
    scalar = rng.standard_normal((n, n, n)).astype(np.float32)
    fields = [(u, (0.5, 0.0, 0.0)), (v, (0.0, 0.5, 0.0)), (w, (0.0, 0.0, 0.5)), (scalar, (0.0, 0.0, 0.0))]
    gu, gv, gw = _t(u), _t(v), _t(w)
    for f, offs in fields:
        ref = O.advect_rk2(h, dt, u, v, w, f, offs, mode=mode)
        out = _t(np.zeros_like(f))
        api.advect_rk2(float(h), float(dt), gu, gv, gw, _t(f), offs, out, mode=mode)
        import torch
        torch.cuda.synchronize()
        assert np.array_equal(bits(_np(out)), bits(ref))


def test_advect_two_fields_share_backtrace(cuda):
    from oracle import oracle as O
    from vflip_amd import api
    import torch
    n = 20
    h = np.float32(1.0 / n)
    dt = np.float32(0.03)
    u, v, w = scenes.random_mac(n, 3)
    rng = np.random.default_rng(0)
    a = rng.random((n, n, n)).astype(np.float32)
    b = rng.random((n, n, n)).astype(np.float32)
    oa, ob = _t(np.zeros_like(a)), _t(np.zeros_like(b))
    api.advect_rk2(float(h), float(dt), _t(u), _t(v), _t(w), _t(a), (0, 0, 0), oa, field2=_t(b), out2=ob)
    torch.cuda.synchronize()
    assert np.array_equal(bits(_np(oa)), bits(O.advect_rk2(h, dt, u, v, w, a, (0, 0, 0))))
    assert np.array_equal(bits(_np(ob)), bits(O.advect_rk2(h, dt, u, v, w, b, (0, 0, 0))))


def test_plume_force_bitexact(cuda):
    from oracle import oracle as O
    from vflip_amd import api
    import torch
    n = 32
    h = np.float32(1.0 / n)
    rng = np.random.default_rng(1)
    rho = rng.random((n, n, n)).astype(np.float32)
    temp = rng.random((n, n, n)).astype(np.float32)
    v = rng.standard_normal((n, n + 1, n)).astype(np.float32)
    args = (float(h), 0.01, (0.5 - 0.5 / n, 0.1 - 0.5 / n, 0.5 - 0.5 / n), 0.08, 0, 0.05, 0.1, 0.3)
    r_ref, t_ref, v_ref = O.plume_force(*args, rho, temp, v)
    gr, gt, gv = _t(rho), _t(temp), _t(v)
    api.plume_force(*args, gr, gt, gv)
    torch.cuda.synchronize()
    for g, r in ((gr, r_ref), (gt, t_ref), (gv, v_ref)):
        assert np.array_equal(bits(_np(g)), bits(r))
    assert (r_ref > 1.0).sum() > 0
