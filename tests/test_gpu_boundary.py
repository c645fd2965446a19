"""GPU parity of the boundary pieces around the projection, against the oracle:

* inflow boundary (S/Smoke_solver3D.cpp:964-980) — bit-exact;
* set_boundary_phi / set_boundary_vel (S/Smoke_solver3D.cpp:793-873): the
  callback sees the reference's node and face positions (oracle grid points,
  pinned by golden vectors) and the device arrays hold its values — bit-exact;
* the assembly with a non-zero solid velocity (S/Smoke_solver3D.cpp:1099-1108,
  S/sub_domain.cpp:721) — within one FP32 rounding of the oracle's FP64 values;
* the SubDomainSolver gradient (dt factor, faces 1..n-1, S/sub_domain.cpp:745-819)
  and the coarse one — bit-exact from the same FP64 pressure;
* the clamped cubic advection on a step field where the cubic overshoots
  (S/Smoke_solver3D.cpp:690-702) — bit-exact.
"""
import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    return t.detach().cpu().numpy()


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def mods():
    from oracle import oracle as O
    from vflip_amd import api
    return O, api


def test_inflow_bc_bitexact(cuda, mods):
    O, api = mods
    ni, nj, nk = 19, 14, 11
    print("inflow bc", flush=True)
    rng = np.random.default_rng(7)
    bd = (rng.random((nk, nj, ni)) < 0.15).astype(np.int8)
    bd[rng.random((nk, nj, ni)) < 0.05] = 2
    u = rng.standard_normal((nk, nj, ni + 1)).astype(np.float32)
    v = rng.standard_normal((nk, nj + 1, ni)).astype(np.float32)
    w = rng.standard_normal((nk + 1, nj, ni)).astype(np.float32)
    ref = O.inflow_bc(bd, np.float32(1.75), u, v, w)
    g = [_t(a) for a in (u, v, w)]
    api.inflow_bc(_t(bd), 1.75, *g)
    for r, x in zip(ref, g):
        assert np.array_equal(_bits(r), _bits(_np(x)))


def test_set_boundary_phi_and_vel_placement(cuda, mods):
    O, api = mods
    ni, nj, nk = 13, 10, 9
    h = np.float32(1.0 / 16)
    print("set_boundary_phi/vel", flush=True)

    def phi(x, y, z):  # sphere SDF in float arithmetic
        c = np.float32(0.31)
        return np.float32(np.sqrt((np.float32(x) - c) ** 2 + (np.float32(y) - np.float32(0.27)) ** 2 +
                                  (np.float32(z) - np.float32(0.22)) ** 2) - np.float32(0.15))

    def vel(x, y, z):  # distinct per component and position
        return (np.float32(x) * np.float32(3) + np.float32(y), np.float32(y) - np.float32(2) * np.float32(z),
                np.float32(z) * np.float32(5) - np.float32(x))

    got = _np(api.set_boundary_phi(ni, nj, nk, float(h), phi))
    pts = O.grid_points((nk + 1, nj + 1, ni + 1), h, (0.5, 0.5, 0.5))
    ref = np.array([phi(*p) for p in pts.reshape(-1, 3)], np.float32).reshape(pts.shape[:3])
    assert np.array_equal(_bits(got), _bits(ref))
    us, vs, ws = api.set_boundary_vel(ni, nj, nk, float(h), vel)
    for comp, (arr, offs) in enumerate(((us, (0.5, 0, 0)), (vs, (0, 0.5, 0)), (ws, (0, 0, 0.5)))):
        g = _np(arr)
        pts = O.grid_points(g.shape, h, offs)
        ref = np.array([vel(*p)[comp] for p in pts.reshape(-1, 3)], np.float32).reshape(g.shape)
        assert np.array_equal(_bits(g), _bits(ref)), comp


@pytest.mark.parametrize("sub", [False, True])
def test_assembly_with_solid_velocity(cuda, mods, sub):
    """Moving solids: random fluid velocity and a non-zero affine solid velocity; the
    GPU operator and right-hand side equal the oracle's FP64 values rounded once."""
    O, api = mods
    n = 28
    print(f"assembly with solid velocity sub={sub}", flush=True)
    pb = scenes.projection_problem(n, 9)
    uw, vw, ww = O.face_weights(pb["phi"])
    us, vs, ws = scenes.affine_mac(n, n, n, float(pb["h"]),
                                   [(0.3, 0.7, -0.2, 0.1), (-0.4, 0.05, -0.9, 0.3), (0.2, 0.4, 0.25, 0.35)])
    dt = np.float32(0.01) if sub else np.float32(1.0)
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], uw, vw, ww, us, vs, ws, dt=dt, sub=sub)
    S = api.Solver(n, n, n)
    rhs_g = _t(np.zeros((n, n, n), np.float32))
    S.assemble(api.SUB if sub else api.COARSE, float(pb["h"]), float(dt),
               *[_t(a) for a in (pb["u"], pb["v"], pb["w"], uw, vw, ww, us, vs, ws)], rhs_g)
    gd, gx, gy, gz = S.get_level(0)
    for ref, got in ((d, gd), (ax, gx), (ay, gy), (az, gz), (rhs, rhs_g)):
        got = _np(got).astype(np.float64)
        assert np.all(np.abs(got - ref) <= 1.2e-7 * np.abs(ref) + 1e-30)
    # the solid term changes the rhs at cut cells (the test exercises it)
    z = scenes.mac_zeros(n)
    *_, rhs0 = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], uw, vw, ww, *z, dt=dt, sub=sub)
    assert np.abs(rhs0 - rhs).max() > 1e-3 * np.abs(rhs).max()


@pytest.mark.parametrize("sub", [False, True])
def test_grad_apply_bitexact(cuda, mods, sub):
    import torch
    O, api = mods
    n = 26
    print(f"grad apply sub={sub}", flush=True)
    pb = scenes.projection_problem(n, 12)
    uw, vw, ww = O.face_weights(pb["phi"])
    us, vs, ws = scenes.affine_mac(n, n, n, float(pb["h"]),
                                   [(0.1, 0.2, 0.0, 0.0), (0.0, 0.0, -0.3, 0.0), (0.05, 0.0, 0.0, 0.4)])
    dt = np.float32(0.02) if sub else np.float32(1.0)
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], uw, vw, ww, us, vs, ws, dt=dt, sub=sub)
    p = O.amgpcg(d, ax, ay, az, rhs, tol=1e-8, maxit=200)["x"]
    ref = O.grad_apply(pb["h"], p, uw, vw, ww, us, vs, ws, pb["u"], pb["v"], pb["w"], dt=dt, sub=sub)
    g = [_t(a) for a in (pb["u"], pb["v"], pb["w"])]
    api.grad_apply_f64(api.SUB if sub else api.COARSE, float(pb["h"]), float(dt),
                       torch.from_numpy(p).cuda(), *[_t(a) for a in (uw, vw, ww, us, vs, ws)], *g)
    for r, x in zip(ref, g):
        assert np.array_equal(_bits(r), _bits(_np(x)))
    # the sub variant also updates the faces next to the ring, the coarse one does not
    changed = np.abs(_np(g[0]) - pb["u"])[1:-1, 1:-1, 1]
    assert (changed.max() > 0) == sub


def test_cubic_clamp_step_field_bitexact(cuda, mods):
    import torch
    O, api = mods
    n = 24
    print("cubic clamp step field", flush=True)
    h = np.float32(1.0 / n)
    dt = np.float32(0.01)
    U = np.float32(-0.25) * h / dt
    u = np.full((n, n, n + 1), U, np.float32)
    v = np.zeros((n, n + 1, n), np.float32)
    w = np.full((n + 1, n, n), np.float32(0.13) * h / dt, np.float32)
    f = np.zeros((n, n, n), np.float32)
    f[:, :, 12:] = 1.0
    f[:, 9:, :] += np.float32(0.5)
    ref = O.advect_rk2(h, dt, u, v, w, f, (0.0, 0.0, 0.0), mode=1)
    lin = O.advect_rk2(h, dt, u, v, w, f, (0.0, 0.0, 0.0), mode=0)
    out = torch.empty((n, n, n), device="cuda")
    api.advect_rk2(float(h), float(dt), _t(u), _t(v), _t(w), _t(f), (0.0, 0.0, 0.0), out, mode=1)
    assert np.array_equal(_bits(ref), _bits(_np(out)))
    assert np.sum(ref != lin) > 0 and np.sum(ref == lin) > 0  # both branches taken
