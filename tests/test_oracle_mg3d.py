"""Pins of the geometric multigrid oracle (MultiGridSolver3D<T>, S/Multigrid3D.h).

The reference class cannot be compiled here (S/Multigrid3D.h includes tbb/tbb.h,
which the image lacks, and the class is never instantiated by the reference, so
it has no tests or vectors). Each oracle function is pinned by an independent
construction instead:
  levels ........... hand-worked extents (m_ComputeLevels' <= 512 rule)
  rbgs ............. matrix form with the 7-point operator built from 1-D second
                     differences (scipy kron), colour order parity 1 then 0; the
                     opposite order and lexicographic Gauss-Seidel both differ;
                     an exact solution is a fixed point
  residual ......... b - L x with the kron operator
  restrict ......... numpy block sums; restrict(L q) = L_2h Q (the 0.5 factor)
                     for q = i^2
  prolongate ....... numpy repeat (injection)
  vcycle / fmg ..... convergence to scipy spsolve (Dirichlet) and to L x = b
                     (Neumann, zero-mean b); a single-level FMG is 200 sweeps
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O


def second_difference(n, dirichlet):
    """1-D part of the operator: neighbours past the end are the cell itself (Neumann)
    or zero (Dirichlet), so the end diagonal is -1 or -2."""
    main = -2.0 * np.ones(n)
    if not dirichlet:
        main[0] += 1.0
        main[-1] += 1.0
    return sp.diags([main, np.ones(n - 1), np.ones(n - 1)], [0, -1, 1], shape=(n, n), format="csr")


def operator(shape, dirichlet):
    """L on x-fastest (nz, ny, nx) arrays: sum of the neighbours - 6 x."""
    nz, ny, nx = shape
    ix, iy, iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    dx, dy, dz = (second_difference(n, dirichlet) for n in (nx, ny, nz))
    return (sp.kron(iz, sp.kron(iy, dx)) + sp.kron(iz, sp.kron(dy, ix)) + sp.kron(dz, sp.kron(iy, ix))).tocsr()


def parity(shape):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i + j + k) % 2).ravel()


def matrix_sweeps(shape, dirichlet, b, x, sweeps, omega, order=(1, 0)):
    """Red-black sweeps in matrix form: per colour, x_c = (1-w) x_c + w ((N x)_c - b_c) / 6
    with N = L + 6 I (same-colour couplings of N are only the diagonal)."""
    L = operator(shape, dirichlet)
    N = L + 6.0 * sp.identity(L.shape[0])
    par = parity(shape)
    x = x.ravel().astype(np.float64).copy()
    bb = b.ravel().astype(np.float64)
    for _ in range(sweeps):
        for colour in order:
            m = par == colour
            y = N @ x
            x[m] = (1 - omega) * x[m] + omega * (y[m] - bb[m]) / 6.0
    return x.reshape(shape)


SHAPES = [(5, 6, 7), (4, 1, 9), (3, 8, 1), (6, 6, 6)]


def test_levels_closed_form():
    assert O.mg3d_levels(256, 256, 256) == [(256, 256, 256), (128, 128, 128), (64, 64, 64), (32, 32, 32),
                                            (16, 16, 16), (8, 8, 8)]
    assert O.mg3d_levels(1024, 16, 16) == [(1024, 16, 16), (512, 8, 8), (256, 4, 4), (128, 2, 2)]
    assert O.mg3d_levels(8, 8, 8) == [(8, 8, 8)]
    assert O.mg3d_levels(9, 8, 8) == [(9, 8, 8), (4, 4, 4)]
    assert O.mg3d_levels(5, 3, 1) == [(5, 3, 1)]
    assert O.mg3d_levels(3, 1, 600) == [(3, 1, 600), (1, 1, 300)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dirichlet", [False, True])
@pytest.mark.parametrize("omega", [1.0, 0.8, 1.3])
def test_rbgs_matrix_form(shape, dirichlet, omega):
    rng = np.random.default_rng(7)
    b = rng.standard_normal(shape)
    x0 = rng.standard_normal(shape)
    got = O.mg3d_rbgs(b, x0, 3, omega, dirichlet)
    want = matrix_sweeps(shape, dirichlet, b, x0, 3, omega)
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
    # the plausible mistakes are visible: red first, or lexicographic order
    wrong = matrix_sweeps(shape, dirichlet, b, x0, 3, omega, order=(0, 1))
    assert np.abs(got - wrong).max() > 1e-3


def test_rbgs_not_lexicographic():
    shape = (4, 5, 6)
    rng = np.random.default_rng(3)
    b, x0 = rng.standard_normal(shape), rng.standard_normal(shape)
    got = O.mg3d_rbgs(b, x0, 1, 1.0, False)
    L = operator(shape, False)
    N = (L + 6.0 * sp.identity(L.shape[0])).tocsr()
    x = x0.ravel().copy()
    for c in range(x.size):  # lexicographic Gauss-Seidel with the same update
        x[c] = ((N[c] @ x)[0] - b.ravel()[c]) / 6.0
    assert np.abs(got.ravel() - x).max() > 1e-3


@pytest.mark.parametrize("dirichlet", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rbgs_fixed_point(dirichlet, dtype):
    shape = (6, 7, 9)
    rng = np.random.default_rng(11)
    x = rng.integers(-20, 20, shape).astype(dtype)
    b = (operator(shape, dirichlet) @ x.ravel().astype(np.float64)).reshape(shape).astype(dtype)
    assert np.array_equal(O.mg3d_rbgs(b, x, 4, 1.0, dirichlet), x)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dirichlet", [False, True])
def test_residual_matrix_form(shape, dirichlet):
    rng = np.random.default_rng(5)
    b, x = rng.standard_normal(shape), rng.standard_normal(shape)
    want = b.ravel() - operator(shape, dirichlet) @ x.ravel()
    assert np.abs(O.mg3d_residual(b, x, dirichlet).ravel() - want).max() <= 1e-12


def block_sum_half(src):
    nz, ny, nx = src.shape
    cz, cy, cx = max(nz // 2, 1), max(ny // 2, 1), max(nx // 2, 1)
    pad = np.zeros((2 * cz, 2 * cy, 2 * cx))
    mz, my, mx = min(nz, 2 * cz), min(ny, 2 * cy), min(nx, 2 * cx)
    pad[:mz, :my, :mx] = src[:mz, :my, :mx]
    return 0.5 * pad.reshape(cz, 2, cy, 2, cx, 2).sum(axis=(1, 3, 5))


@pytest.mark.parametrize("shape", [(8, 6, 4), (7, 5, 9), (1, 4, 6), (3, 1, 1)])
def test_restrict_block_sum(shape):
    src = np.random.default_rng(2).standard_normal(shape)
    got = O.mg3d_restrict(src)
    assert np.abs(got - block_sum_half(src)).max() <= 1e-14 * 8


def test_restrict_is_coarse_operator():
    # q = i^2 (fine index): L q = 2 at interior cells; the coarse values Q(I) = (2I + 1/2)^2
    # have L_2h Q = 8, which the restricted residual must reproduce.
    shape = (8, 8, 16)
    i = np.arange(shape[2], dtype=np.float64)
    q = np.broadcast_to(i * i, shape).copy()
    r = O.mg3d_restrict(O.mg3d_residual(np.zeros(shape), q))  # = restrict(-L q)
    assert np.array_equal(r[1:-1, 1:-1, 1:-1], np.full((2, 2, 6), -8.0))
    I = np.arange(shape[2] // 2, dtype=np.float64)
    Q = np.broadcast_to((2 * I + 0.5) ** 2, tuple(n // 2 for n in shape)).copy()
    LQ = -O.mg3d_residual(np.zeros(Q.shape), Q)
    assert np.array_equal(LQ[1:-1, 1:-1, 1:-1], r[1:-1, 1:-1, 1:-1] * -1.0)


@pytest.mark.parametrize("shape", [(8, 6, 4), (7, 5, 9), (1, 4, 6), (3, 1, 1)])
def test_prolongate_injection(shape):
    rng = np.random.default_rng(4)
    dst = rng.standard_normal(shape)
    cshape = tuple(max(n // 2, 1) for n in shape)
    src = rng.standard_normal(cshape)
    rep = src.repeat(2, 0).repeat(2, 1).repeat(2, 2)
    add = np.zeros(shape)
    mz, my, mx = (min(a, b) for a, b in zip(shape, rep.shape))
    add[:mz, :my, :mx] = rep[:mz, :my, :mx]
    assert np.array_equal(O.mg3d_prolongate(src, dst), dst + add)


def test_vcycle_dirichlet_converges_to_direct_solve():
    shape = (24, 20, 32)
    b = np.random.default_rng(9).standard_normal(shape)
    want = spla.spsolve(operator(shape, True).tocsc(), b.ravel()).reshape(shape)
    x = np.zeros(shape)
    norms = []
    for _ in range(10):
        x = O.mg3d_solve(b, x, 1, dirichlet=True)
        norms.append(np.abs(O.mg3d_residual(b, x, True)).max())
    assert all(n1 < 0.2 * n0 for n0, n1 in zip(norms, norms[1:]))
    assert np.abs(x - want).max() <= 1e-8 * np.abs(want).max()


def test_vcycle_neumann_solves_zero_mean():
    shape = (16, 16, 16)
    b = np.random.default_rng(10).standard_normal(shape)
    b -= b.mean()
    x = O.mg3d_solve(b, np.zeros(shape), 12)
    r = b.ravel() - operator(shape, False) @ x.ravel()
    assert np.abs(r).max() <= 1e-7 * np.abs(b).max()


def test_fmg_single_level_is_200_sweeps():
    shape = (6, 7, 8)
    rng = np.random.default_rng(12)
    b, x0 = rng.standard_normal(shape), rng.standard_normal(shape)
    assert np.array_equal(O.mg3d_solve(b, x0, 0, fmg=True, dirichlet=True), O.mg3d_rbgs(b, x0, 200, 1.0, True))
    assert np.array_equal(O.mg3d_solve(b, x0, 1, dirichlet=True), O.mg3d_rbgs(b, x0, 100, 1.0, True))


def test_fmg_beats_one_vcycle_and_converges():
    shape = (32, 32, 32)
    b = np.random.default_rng(13).standard_normal(shape)
    want = spla.spsolve(operator(shape, True).tocsc(), b.ravel()).reshape(shape)
    xf = O.mg3d_solve(b, np.zeros(shape), 0, fmg=True, dirichlet=True)
    xv = O.mg3d_solve(b, np.zeros(shape), 1, dirichlet=True)
    ef, ev = np.abs(xf - want).max(), np.abs(xv - want).max()
    assert ef < 0.5 * ev
    x = O.mg3d_solve(b, xf, 8, dirichlet=True)
    assert np.abs(x - want).max() <= 1e-8 * np.abs(want).max()


def test_float_tracks_double():
    shape = (16, 12, 20)
    b = np.random.default_rng(14).standard_normal(shape)
    xd = O.mg3d_solve(b, np.zeros(shape), 3, dirichlet=True)
    xf = O.mg3d_solve(b.astype(np.float32), np.zeros(shape, np.float32), 3, dirichlet=True)
    assert xf.dtype == np.float32
    assert np.abs(xf - xd).max() <= 1e-5 * np.abs(xd).max()
