"""The C++ drop-in facade (include/vflip_ivock.hpp, vflip_amd/libvflip_ivock.so) driven
the way the reference's code drives SmokeSolver3D and AMGPCGSolve, by the compiled
check program vflip_amd/facade/facade_check (built with the library):

* set_boundary_phi / compute_face_weights / pcg_projection on a 24^3 grid with a
  sphere and box walls: face weights bit-exact with the oracle, the right-hand side
  within one FP32 rounding, the FP64 true residual at the tolerance and the projected
  velocity within 1e-5 relative L2 of the oracle's FP64 projection;
* the FLIP particles init() seeds, hash_FLIP and particle_to_grid for u, v, w:
  bit-exact with the oracle's gather in the reference's summation order;
* ivock::AMGPCGSolve<double> (the shim of INTEGRATION.md, in the header) on a
  SparseMatrixd cut-cell system: converged, iterations within one of the oracle's,
  residual_out equal to the FP64 |b - A x|_inf of the result and below 1e-6 |b|_inf;
* MGSolverf / MGSolverd (MultiGridSolver3D): level operators, V-cycles and full
  multigrid bit-exact with the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "vflip_amd", "facade", "facade_check")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def run(tmp_path_factory, cuda):
    out = tmp_path_factory.mktemp("facade")
    print("facade check program", flush=True)
    r = subprocess.run([CHECK, str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    n = 24

    def load(name, shape=None):
        dt = np.float64 if name.endswith("f64") else np.float32
        a = np.fromfile(os.path.join(out, name), dtype=dt)
        return a.reshape(shape) if shape else a
    return load, n


def test_facade_projection(run):
    from oracle import oracle as O
    load, n = run
    h = np.float32(1.0 / n)
    mac = {"u": (n, n, n + 1), "v": (n, n + 1, n), "w": (n + 1, n, n)}
    phi = load("phi.f32", (n + 1, n + 1, n + 1))
    uw, vw, ww = O.face_weights(phi)
    for ref, name in ((uw, "uw"), (vw, "vw"), (ww, "ww")):
        assert np.array_equal(_bits(ref), _bits(load(name + ".f32", ref.shape)))
    u0, v0, w0 = (load(k + "0.f32", mac[k]) for k in "uvw")
    z = scenes.mac_zeros(n)
    d, ax, ay, az, rhs = O.assemble(h, u0, v0, w0, uw, vw, ww, *z)
    got_rhs = load("rhs.f64", (n, n, n))
    assert np.all(np.abs(got_rhs - rhs) <= 1.2e-7 * np.abs(rhs) + 1e-30)
    its, rel_res, true_res = load("stats.f64")
    ref = O.amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50)
    print(f"facade projection: {int(its)} its (oracle {ref['iterations']}), true residual {true_res:.2e}", flush=True)
    assert true_res <= 1e-6 and abs(int(its) - ref["iterations"]) <= 1
    ru, rv, rw = O.grad_apply(h, ref["x"], uw, vw, ww, *z, u0, v0, w0)
    for r, k in ((ru, "u"), (rv, "v"), (rw, "w")):
        assert _rel(load(k + "1.f32", mac[k]), r) < 1e-5


def test_facade_particle_to_grid_bitexact(run):
    from oracle import oracle as O
    load, n = run
    h = np.float32(1.0 / n)
    pos = load("fpos.f32").reshape(-1, 3)
    vel = load("fvel.f32").reshape(-1, 3)
    assert len(pos) > 0
    _, start, order = O.bin_particles(pos, h, (n, n, n))
    specs = ((0, (n, n, n + 1), (0.5, 0, 0), "u"), (1, (n, n + 1, n), (0, 0.5, 0), "v"),
             (2, (n + 1, n, n), (0, 0, 0.5), "w"))
    for comp, shape, offs, k in specs:
        f, c = O.p2g(pos, vel, comp, h, (n, n, n), start, order, shape, offs)
        assert np.array_equal(_bits(f), _bits(load("p" + k + ".f32", shape))), k
        assert np.array_equal(_bits(c), _bits(load("c" + k + ".f32", shape))), k


def test_facade_amgpcg_shim(run):
    from oracle import oracle as O
    load, _ = run
    ni, nj, nk = 20, 18, 16
    shp = (nk, nj, ni)
    d, ax, ay, az, b, x = (load(f"amg_{k}.f64", shp) for k in ("d", "ax", "ay", "az", "b", "x"))
    ok, its, res = load("amg_stats.f64")
    ref = O.amgpcg(d, ax, ay, az, b, tol=1e-6, maxit=50)
    true_res = np.abs(b - O.matvec(d, ax, ay, az, x)).max()
    print(f"AMGPCGSolve shim: {int(its)} its (oracle {ref['iterations']}), residual_out {res:.3e}, "
          f"true {true_res:.3e}", flush=True)
    assert ok == 1.0 and abs(int(its) - ref["iterations"]) <= 1
    assert true_res <= 1e-6 * np.abs(b).max()
    assert abs(res - true_res) <= 1e-9 * np.abs(b).max()
    assert _rel(x, ref["x"]) < 1e-4


def test_facade_mg3d_bitexact(run):
    """MGSolverf / MGSolverd through the facade: residual, restriction, two V-cycles and
    full multigrid with Dirichlet walls in float, a Neumann V-cycle in double; all
    bit-exact with the oracle."""
    from oracle import oracle as O
    load, _ = run
    shape = (24, 18, 20)
    b = load("mg_b.f32", shape)
    x0 = load("mg_x0.f32", shape)
    assert np.array_equal(_bits(load("mg_res.f32", shape)), _bits(O.mg3d_residual(b, x0, True)))
    assert np.array_equal(_bits(load("mg_restrict.f32", (12, 9, 10))), _bits(O.mg3d_restrict(O.mg3d_residual(b, x0, True))))
    assert np.array_equal(_bits(load("mg_vcycle2.f32", shape)), _bits(O.mg3d_solve(b, x0, 2, dirichlet=True)))
    assert np.array_equal(_bits(load("mg_fmg.f32", shape)), _bits(O.mg3d_solve(b, x0, 0, fmg=True, dirichlet=True)))
    want = O.mg3d_solve(b.astype(np.float64), x0.astype(np.float64), 1, dirichlet=False)
    got = load("mg_vcycle_neumann.f64", shape)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
