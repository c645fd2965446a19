"""GPU advection and plume forcing against the oracle. Modes 0/1 and the forcing
are bit-exact (same FP64 interpolation, same float/double rounding points, no FMA
contraction); mode 2 (the same trilinear RK2 in FP32 with fused lerps, used by
the simulation step) is within 1e-6 relative L2 of the oracle's FP64 result and
within what a 1-ulp difference of a float backtrace position changes, cell by
cell."""
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


@pytest.mark.parametrize("n", [24, 33, 64])
def test_advect_fp32_mode_close_to_oracle(cuda, n):
    from oracle import oracle as O
    from vflip_amd import api
    import torch
    print(f"advect fp32 n={n}", flush=True)
    h = np.float32(1.0 / n)  # n = 24, 33: general first RK2 stage; n = 64 (power-of-two h): lattice form
    dt = np.float32(0.6 / n)
    u, v, w = scenes.random_mac(n, 23, scale=1.0)
    rng = np.random.default_rng(n + 1)
    a = rng.standard_normal((n, n, n)).astype(np.float32)
    b = rng.random((n, n, n)).astype(np.float32)
    gu, gv, gw = _t(u), _t(v), _t(w)
    for f, offs, f2 in ((u, (0.5, 0.0, 0.0), None), (v, (0.0, 0.5, 0.0), None), (w, (0.0, 0.0, 0.5), None),
                        (a, (0.0, 0.0, 0.0), b)):
        out = _t(np.zeros_like(f))
        out2 = _t(np.zeros_like(f2)) if f2 is not None else None
        api.advect_rk2(float(h), float(dt), gu, gv, gw, _t(f), offs, out, mode=2,
                       field2=_t(f2) if f2 is not None else None, out2=out2)
        torch.cuda.synchronize()
        for got, src in ((out, f), (out2, f2)):
            if got is None:
                continue
            ref = O.advect_rk2(h, dt, u, v, w, src, offs, mode=0).astype(np.float64)
            g = _np(got).astype(np.float64)
            assert np.linalg.norm(g - ref) <= 1e-6 * np.linalg.norm(ref)
            # per cell: a 1-ulp difference of a float position in [0, 1] (<= 1.2e-7) moves
            # the sample by <= 1.2e-7 * n cells; on these rough fields (values change by up
            # to 2 max|f| per cell) that bounds the cell error by 2.4e-7 * n * max|f|
            assert np.abs(g - ref).max() <= 2.5e-7 * n * np.abs(src).max()


@pytest.mark.parametrize("n,h", [(40, None), (64, None), (48, 1.0 / 48)])
def test_fused_plume_advection_bitexact(cuda, n, h):
    """vflip_advect_plume equals the five mode-2 vflip_advect_rk2 calls bit for bit: with a
    power-of-two h (lattice first stage) and without, velocities large enough that some
    backtraces leave the staged tile (global fallback), ragged tile edges."""
    import torch
    from vflip_amd import api
    print(f"fused plume advection n={n}", flush=True)
    nk, nj, ni = n - 3, n - 6, n  # ragged against the 32 x 8 x 4 tiles
    hh = 1.0 / 64 if h is None else h
    g = torch.Generator(device="cuda").manual_seed(n)
    mk = lambda *s: torch.rand(s, device="cuda", generator=g) * 2 - 1
    u, v, w = mk(nk, nj, ni + 1) * 2.0, mk(nk, nj + 1, ni) * 2.0, mk(nk + 1, nj, ni) * 2.0
    rho, temp = mk(nk, nj, ni), mk(nk, nj, ni)
    dt = 0.02  # |u| dt / h up to ~2.6 cells: beyond the 2-cell halo
    ref = [torch.empty_like(a) for a in (u, v, w, rho, temp)]
    api.advect_rk2(hh, dt, u, v, w, u, (0.5, 0, 0), ref[0], mode=2)
    api.advect_rk2(hh, dt, u, v, w, v, (0, 0.5, 0), ref[1], mode=2)
    api.advect_rk2(hh, dt, u, v, w, w, (0, 0, 0.5), ref[2], mode=2)
    api.advect_rk2(hh, dt, u, v, w, rho, (0, 0, 0), ref[3], mode=2, field2=temp, out2=ref[4])
    got = [torch.empty_like(a) for a in (u, v, w, rho, temp)]
    api.advect_plume(hh, dt, u, v, w, rho, temp, *got)
    torch.cuda.synchronize()
    for a, b in zip(ref, got):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
