"""GPU parity of the pressure projection (HIP, via the C ABI) against the oracle.

Bars: face weights bit-exact; assembled operator within FP32 rounding of the
oracle's FP64 values; one V-cycle and the full PCG solve within 1e-5 relative
L2 of the oracle (same iteration count, +-1 for FP32 vs FP64 round-off), and the
returned pressure's FP64 true residual at the requested tolerance level.
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


def oracle_iterate(O, d, ax, ay, az, rhs, iterations):
    """The oracle's PCG iterate after exactly `iterations` iterations (tol 0)."""
    return O.amgpcg(d, ax, ay, az, rhs, tol=0.0, maxit=iterations)["x"]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def mods():
    from oracle import oracle as O
    from vflip_amd import api
    return O, api


def build_problem(n, seed, O):
    pb = scenes.projection_problem(n, seed)
    uw, vw, ww = O.face_weights(pb["phi"])
    pb.update(uw=uw, vw=vw, ww=ww)
    return pb


@pytest.mark.parametrize("n,seed", [(24, 1), (33, 2), (64, 3)])
def test_face_weights_bitexact(cuda, mods, n, seed):
    O, api = mods
    pb = scenes.projection_problem(n, seed)
    print(f"face weights n={n} seed={seed}", flush=True)
    ref = O.face_weights(pb["phi"])
    got = api.face_weights(_t(pb["phi"]))
    for r, g in zip(ref, got):
        assert np.array_equal(r.view(np.uint32), _np(g).view(np.uint32))


@pytest.mark.parametrize("n,sub", [(32, False), (31, True)])
def test_assembly(cuda, mods, n, sub):
    O, api = mods
    pb = build_problem(n, 5, O)
    dt = np.float32(0.01) if sub else np.float32(1.0)
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"], dt=dt, sub=sub)
    S = api.Solver(n, n, n)
    rhs_g = _t(np.zeros((n, n, n), np.float32))
    args = [_t(pb[k]) for k in ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")]
    S.assemble(api.SUB if sub else api.COARSE, float(pb["h"]), float(dt), *args, rhs_g)
    gd, gx, gy, gz = S.get_level(0)
    for ref, got in ((d, gd), (ax, gx), (ay, gy), (az, gz), (rhs, rhs_g)):
        got = _np(got).astype(np.float64)
        # FP64 accumulation rounded once to FP32
        assert np.all(np.abs(got - ref) <= 1.2e-7 * np.abs(ref) + 1e-30)


@pytest.mark.parametrize("n", [40, 64])
def test_coarse_levels(cuda, mods, n):
    O, api = mods
    pb = build_problem(n, 7, O)
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"])
    S = api.Solver(n, n, n)
    S.set_matrix(*[_t(a.astype(np.float32)) for a in (d, ax, ay, az)])
    S.setup()
    levels = S.levels()
    assert levels == O.level_dims(n, n, n)
    cur = (d, ax, ay, az)
    for l in range(1, len(levels)):
        cur = O.coarsen(*cur)
        got = S.get_level(l)
        for ref, g in zip(cur, got):
            g = _np(g).astype(np.float64)
            scale = np.abs(cur[0]).max()
            assert np.abs(g - ref).max() <= 2e-6 * scale


def _solver_with_matrix(api, O, pb, n, **opts):
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"])
    S = api.Solver(n, n, n, **opts)
    S.set_matrix(*[_t(a.astype(np.float32)) for a in (d, ax, ay, az)])
    return S, (d, ax, ay, az, rhs)


@pytest.mark.parametrize("n", [32, 48])
def test_vcycle(cuda, mods, n):
    O, api = mods
    pb = build_problem(n, 11, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n)
    z_ref = O.vcycle(d, ax, ay, az, rhs)
    z = _t(np.zeros((n, n, n), np.float32))
    S.vcycle(_t(rhs.astype(np.float32)), z)
    assert rel_l2(_np(z), z_ref) < 1e-5


@pytest.mark.parametrize("n,seed", [(32, 1), (64, 2), (50, 3)])
def test_pcg_parity(cuda, mods, n, seed):
    O, api = mods
    print(f"pcg parity n={n} seed={seed}", flush=True)
    pb = build_problem(n, seed, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n)
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    x = _t(np.zeros((n, n, n), np.float32))
    st = S.solve(_t(rhs.astype(np.float32)), x)
    assert ref["ok"] and st["status"] == 0
    assert abs(st["iterations"] - ref["iterations"]) <= 1
    # same algorithm, same iteration count -> same iterate up to FP32 round-off
    assert rel_l2(_np(x), oracle_iterate(O, d, ax, ay, az, rhs, st["iterations"])) < 1e-5
    # FP64 residual of the GPU answer against the FP64 operator
    r = rhs - O.matvec(d, ax, ay, az, _np(x).astype(np.float64))
    assert np.abs(r).max() <= 5e-6 * np.abs(rhs).max()


def test_pcg_mixed(cuda, mods):
    O, api = mods
    n = 64
    pb = build_problem(n, 4, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n, precision=api.MIXED)
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    x = _t(np.zeros((n, n, n), np.float32))
    st = S.solve(_t(rhs.astype(np.float32)), x)
    assert st["status"] == 0
    assert st["iterations"] <= ref["iterations"] + 2
    assert st["true_residual"] < 5e-6
    assert rel_l2(_np(x), oracle_iterate(O, d, ax, ay, az, rhs, st["iterations"])) < 1e-5


def test_zero_rhs_and_nonconvergence(cuda, mods):
    O, api = mods
    n = 24
    pb = build_problem(n, 9, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n)
    x = _t(np.ones((n, n, n), np.float32))
    st = S.solve(_t(np.zeros((n, n, n), np.float32)), x)
    assert st["status"] == 0 and st["iterations"] == 0 and float(x.abs().max()) == 0.0
    S.set_opts(max_iterations=1)
    st = S.solve(_t(rhs.astype(np.float32)), x)
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=1)
    assert st["status"] == 1 and not ref["ok"] and st["iterations"] == 1


def test_project_matches_oracle(cuda, mods):
    O, api = mods
    n = 48
    pb = build_problem(n, 13, O)
    h = float(pb["h"])
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"])
    ref = O.amgpcg(d, ax, ay, az, rhs)
    S = api.Solver(n, n, n)
    g = {k: _t(pb[k]) for k in ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")}
    rhs_g = _t(np.zeros((n, n, n), np.float32))
    p = _t(np.zeros((n, n, n), np.float32))
    st = S.project(api.COARSE, h, 1.0, g["u"], g["v"], g["w"], g["uw"], g["vw"], g["ww"], g["us"], g["vs"],
                   g["ws"], rhs_g, p)
    assert st["status"] == 0 and abs(st["iterations"] - ref["iterations"]) <= 1
    # The oracle solves the very FP32 operator and right-hand side the GPU
    # assembled (assembly parity is test_assembly): rounding the operator to FP32
    # alone moves this scene's FP64 solution by ~3e-5 (a near-singular cut-cell
    # pocket), which no FP32 solver can undo. The error against the unrounded
    # FP64 solve is reported, not asserted.
    gd, gx, gy, gz = (_np(a).astype(np.float64) for a in S.get_level(0))
    p_ref = oracle_iterate(O, gd, gx, gy, gz, _np(rhs_g).astype(np.float64), st["iterations"])
    assert rel_l2(_np(p), p_ref) < 1e-5
    print(f"error vs unrounded FP64 oracle: {rel_l2(_np(p), ref['x']):.2e}")
    ru, rv, rw = O.grad_apply(pb["h"], p_ref, pb["uw"], pb["vw"], pb["ww"], pb["us"], pb["vs"], pb["ws"],
                              pb["u"], pb["v"], pb["w"])
    for a, b in ((g["u"], ru), (g["v"], rv), (g["w"], rw)):
        assert rel_l2(_np(a), b) < 1e-5


def test_host_dropin(cuda, mods):
    """vflip_amgpcg_solve_host, the AMGPCGSolve<double> drop-in. Default (no options):
    the mixed solver, so residual_out is the FP64 |b - A x|_inf of the returned result,
    as the reference's; explicit FP32: the plain solver's iterate equals the oracle's."""
    O, api = mods
    n = 32
    pb = build_problem(n, 21, O)
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"])
    ref = O.amgpcg(d, ax, ay, az, rhs)
    rc, x, res, its = api.amgpcg_solve_host(d, ax, ay, az, rhs)
    true_res = np.abs(rhs - O.matvec(d, ax, ay, az, x)).max()
    assert rc == 0 and abs(its - ref["iterations"]) <= 1
    assert res <= 1e-6 * np.abs(rhs).max() and true_res <= 1e-6 * np.abs(rhs).max()
    assert abs(res - true_res) <= 1e-9 * np.abs(rhs).max()
    rc, x, res, its = api.amgpcg_solve_host(d, ax, ay, az, rhs, precision=api.FP32)
    assert rc == 0 and abs(its - ref["iterations"]) <= 1
    assert res <= 1e-6 * np.abs(rhs).max()
    assert rel_l2(x, oracle_iterate(O, d, ax, ay, az, rhs, its)) < 1e-5


@pytest.mark.parametrize("dims", [(30, 34, 32), (9, 129, 130)])
def test_coarsest_beyond_register_capacity(cuda, mods, dims):
    """Grids whose coarsest level has more x-pairs than the register-resident
    coarsest kernel holds (15 x 17 x 16: 2176 pairs; 5 x 65 x 65: 12675 > 4096
    unknowns is never coarsest, 9 x 129 x 130 ends at 3 x 33 x 33 with 2178 pairs):
    the general coarsest kernel solves them, in the V-cycle and in the tabulated
    form, and the solve matches the oracle."""
    import torch
    O, api = mods
    ni, nj, nk = dims
    print(f"coarsest beyond register capacity {dims}", flush=True)
    pb = scenes.box_problem(ni, nj, nk, 4)
    uw, vw, ww = O.face_weights(pb["phi"])
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], uw, vw, ww, pb["us"], pb["vs"], pb["ws"])
    S = api.Solver(ni, nj, nk)
    lv = S.levels()[-1]
    assert ((lv[0] + 1) // 2) * lv[1] * lv[2] > 2048 and lv[0] * lv[1] * lv[2] <= 4096
    S.set_matrix(*[_t(a.astype(np.float32)) for a in (d, ax, ay, az)])
    x = torch.zeros((nk, nj, ni), device="cuda")
    for rep in range(2):  # the first solve sweeps directly, a reused operator is tabulated
        x.zero_()
        st = S.solve(_t(rhs.astype(np.float32)), x)
        ref = O.amgpcg(d, ax, ay, az, rhs, tol=0.0, maxit=st["iterations"])
        print(rep, st["iterations"], flush=True)
        assert st["status"] == 0
        assert rel_l2(_np(x), ref["x"]) < 1e-5
        S.set_matrix(*[_t(a.astype(np.float32)) for a in (d, ax, ay, az)])


def test_repeated_assembly_before_solve(cuda, mods):
    """assemble(A); solve; assemble(B); assemble(B); solve: the change flag set by the
    first B assembly survives the second (which changes no bits), so the second solve
    rebuilds the hierarchy and matches a fresh solver. Mixed precision, whose V-cycle
    reads FP16 copies of every level including level 0: a stale hierarchy would
    precondition with the first operator altogether."""
    import torch
    O, api = mods
    n = 40
    print("repeated assembly before solve", flush=True)
    S = api.Solver(n, n, n, precision=api.MIXED)
    res = []
    for seeds in ((1,), (2, 2)):
        for seed in seeds:
            pb = build_problem(n, seed, O)
            args = [_t(pb[k]) for k in ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")]
            rhs = torch.zeros((n, n, n), device="cuda")
            S.assemble(api.COARSE, float(pb["h"]), 1.0, *args, rhs)
        x = torch.zeros_like(rhs)
        st = S.solve(rhs, x)
        rf = torch.zeros_like(rhs)
        G = api.Solver(n, n, n, precision=api.MIXED)
        G.assemble(api.COARSE, float(pb["h"]), 1.0, *args, rf)
        xf = torch.zeros_like(rhs)
        sf = G.solve(rf, xf)
        print(seeds, st["iterations"], sf["iterations"], st["true_residual"], flush=True)
        assert st["converged"] == 1 and abs(st["iterations"] - sf["iterations"]) <= 1
        assert st["true_residual"] <= 1e-6
        assert rel_l2(_np(x), _np(xf)) < 1e-5
        res.append(st["iterations"])


@pytest.mark.parametrize("n,seed", [(64, 2), (96, 8)])
def test_mixed_reaches_fp64_true_residual(cuda, mods, n, seed):
    """Mixed mode: FP64 iterate + residual replacement reach |b - A x|_inf <= tol |b|_inf
    against the exact FP64 operator and right-hand side."""
    import torch
    O, api = mods
    pb = build_problem(n, seed, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n, precision=api.MIXED)
    S.set_diag64(_t(d))
    x = torch.zeros((n, n, n), device="cuda", dtype=torch.float64)
    st = S.solve_f64(_t(rhs), x)
    assert st["status"] == 0
    xr = _np(x)
    r = rhs - O.matvec(d, ax, ay, az, xr)
    assert np.abs(r).max() <= 1.0e-6 * np.abs(rhs).max()
    assert st["true_residual"] <= 1.0e-6
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    exact = O.amgpcg(d, ax, ay, az, rhs, tol=1e-12, maxit=200)["x"]
    e_gpu, e_ref = rel_l2(xr, exact), rel_l2(ref["x"], exact)

    def energy(e):  # A-norm, the norm CG minimises
        return float(np.sqrt(max(np.sum(e * O.matvec(d, ax, ay, az, e)), 0.0)))

    ea_gpu = energy(xr - exact) / energy(exact)
    ea_ref = energy(ref["x"] - exact) / energy(exact)
    print(f"mixed n={n}: its {st['iterations']} (FP64 oracle {ref['iterations']}), "
          f"true residual {st['true_residual']:.2e}, L2 error vs exact {e_gpu:.2e} (FP64 oracle at 1e-6: "
          f"{e_ref:.2e}), A-norm error {ea_gpu:.2e} (oracle {ea_ref:.2e})")
    # Two solutions at |r|_inf <= 1e-6 |b|_inf can differ by up to cond(A)*1e-6 in L2
    # (near-singular cut-cell pockets); same iteration count and the same A-norm
    # accuracy class are required, the L2 error is reported.
    assert abs(st["iterations"] - ref["iterations"]) <= 1
    assert ea_gpu <= 10 * ea_ref + 1e-7


def test_project_mixed_fp64_pressure(cuda, mods):
    O, api = mods
    n = 48
    pb = build_problem(n, 13, O)
    h = float(pb["h"])
    d, ax, ay, az, rhs = O.assemble(pb["h"], pb["u"], pb["v"], pb["w"], pb["uw"], pb["vw"], pb["ww"],
                                    pb["us"], pb["vs"], pb["ws"])
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6)
    S = api.Solver(n, n, n, precision=api.MIXED)
    g = {k: _t(pb[k]) for k in ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")}
    rhs_g = _t(np.zeros((n, n, n), np.float32))
    p = _t(np.zeros((n, n, n), np.float32))
    st = S.project(api.COARSE, h, 1.0, g["u"], g["v"], g["w"], g["uw"], g["vw"], g["ww"], g["us"], g["vs"],
                   g["ws"], rhs_g, p)
    assert st["status"] == 0 and st["true_residual"] <= 1e-6
    # The FP64 system is the oracle's (same assembly in FP64). This scene has a
    # near-singular cut-cell pocket, so two 1e-6 solutions differ in a nearly
    # constant pocket pressure (reported); the A-norm error and the projected
    # velocities, which only see pressure differences, are asserted.
    pg = _np(p).astype(np.float64)
    print(f"project mixed: p rel L2 vs oracle {rel_l2(pg, ref['x']):.2e}")
    e = pg - ref["x"]
    assert np.sqrt(abs(np.sum(e * O.matvec(d, ax, ay, az, e)))) <= 1e-5 * np.sqrt(
        np.sum(ref["x"] * O.matvec(d, ax, ay, az, ref["x"])))
    ru, rv, rw = O.grad_apply(pb["h"], ref["x"], pb["uw"], pb["vw"], pb["ww"], pb["us"], pb["vs"], pb["ws"],
                              pb["u"], pb["v"], pb["w"])
    for a, b in ((g["u"], ru), (g["v"], rv), (g["w"], rw)):
        assert rel_l2(_np(a), b) < 1e-5


def test_reassembly_rebuilds_hierarchy(cuda, mods):
    """vflip_solver_solve after vflip_assemble with a different operator rebuilds the
    multigrid hierarchy: the reused solver converges like a fresh one (same iterations
    +-1, solutions within FP32 round-off), and an unchanged re-assembly keeps it.
    (Before this was handled, a solver re-assembled with a new liquid region kept
    its first hierarchy and stalled at the iteration limit.)"""
    print("reassembly", flush=True)
    import torch
    O, api = mods
    n = 40
    S = api.Solver(n, n, n)
    outs = []
    for seed in (1, 2, 2):
        pb = build_problem(n, seed, O)
        args = [_t(pb[k]) for k in ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")]
        rhs = torch.zeros((n, n, n), device="cuda")
        S.assemble(api.COARSE, float(pb["h"]), 1.0, *args, rhs)
        x = torch.zeros_like(rhs)
        st = S.solve(rhs, x)
        F = api.Solver(n, n, n)
        rf = torch.zeros_like(rhs)
        F.assemble(api.COARSE, float(pb["h"]), 1.0, *args, rf)
        xf = torch.zeros_like(rhs)
        sf = F.solve(rf, xf)
        assert st["converged"] == 1 and abs(st["iterations"] - sf["iterations"]) <= 1
        assert rel_l2(_np(x), _np(xf)) < 1e-5
        outs.append(_np(x))
    # same operator again: the hierarchy is kept (the coarsest level is now tabulated
    # instead of swept directly, equal up to FP32 rounding)
    assert rel_l2(outs[2], outs[1]) < 1e-6


def test_mixed_forms_and_switching(cuda, mods):
    """The mixed solver's two forms, FP16 coefficient copies of every level with a refresh
    every 4 iterations (the default) and of levels >= 2 with a refresh every 8 (the form the
    plume configuration words), both stop on an FP64 true residual within the tolerance. The
    solver is switched between them and back on one handle, which rebuilds the copies each
    time. The default form needs no more iterations than the other; the FP64 oracle's
    count is the reference."""
    import torch
    O, api = mods
    n = 64
    pb = build_problem(n, 6, O)
    S, (d, ax, ay, az, rhs) = _solver_with_matrix(api, O, pb, n, precision=api.MIXED)
    S.set_diag64(_t(d))
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    dflt = api.default_opts()
    assert (dflt.half_from_level, dflt.refine_every) == (0, 4)
    its = {}
    for form, opts in (("default", {}), ("configuration", {"half_from_level": 2, "refine_every": 8}),
                       ("default again", {"half_from_level": 0, "refine_every": 4})):
        if opts:
            S.set_opts(**opts)
        x = torch.zeros((n, n, n), device="cuda", dtype=torch.float64)
        st = S.solve_f64(_t(rhs), x)
        r = rhs - O.matvec(d, ax, ay, az, _np(x))
        print(f"{form}: its {st['iterations']} (oracle {ref['iterations']}), true residual {st['true_residual']:.2e}",
              flush=True)
        assert st["status"] == 0 and st["true_residual"] <= 1e-6
        assert np.abs(r).max() <= 1.0e-6 * np.abs(rhs).max()
        its[form] = st["iterations"]
    assert its["default"] == its["default again"]
    assert its["default"] <= its["configuration"]
    assert its["default"] <= ref["iterations"] + 1
