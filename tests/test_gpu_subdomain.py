"""GPU coarse/fine coupling operators and brick storage against the oracle:
bit-exact (same loop order, FP64 samplers, no FMA contraction)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# (coarse cells, ratio): the reference's 8x nest and the 2x nests of the nested workload
CASES = [((6, 5, 4), 8), ((12, 10, 9), 2), ((7, 7, 7), 4)]


@pytest.mark.parametrize("cdims,N", CASES)
def test_fine_to_coarse_bitexact(cuda, cdims, N):
    print(f"fine_to_coarse {cdims} N={N}", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    cx, cy, cz = cdims
    rng = np.random.default_rng(N)
    # cell-centred scalar (offsets 0) and the three MAC components (offset N/2 along the normal)
    for axis in (None, 0, 1, 2):
        fshape = [cz * N, cy * N, cx * N]
        cshape = [cz, cy, cx]
        offs = [0, 0, 0]
        if axis is not None:
            fshape[2 - axis] += 1
            cshape[2 - axis] += 1
            offs[axis] = N // 2
        fine = rng.standard_normal(fshape).astype(np.float32)
        ref = np.zeros(cshape, np.float32)
        O.lib().or_fine_to_coarse(fine, fshape[2], fshape[1], fshape[0], ref, cshape[2], cshape[1], cshape[0], N,
                                  *offs)
        got = _t(np.zeros(cshape, np.float32))
        api.fine_to_coarse(_t(fine), got, N, offs)
        assert np.array_equal(bits(_np(got)), bits(ref))


@pytest.mark.parametrize("cdims,N", CASES)
def test_coarse_to_fine_and_small_difference_bitexact(cuda, cdims, N):
    print(f"coarse_to_fine {cdims} N={N}", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    cx, cy, cz = cdims
    hc = np.float32(1.0 / cx)
    hf = np.float32(hc / N)
    rng = np.random.default_rng(10 + N)
    for offs in ((0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5)):
        cshape = (cz + (offs[2] > 0), cy + (offs[1] > 0), cx + (offs[0] > 0))
        fshape = (cz * N + (offs[2] > 0), cy * N + (offs[1] > 0), cx * N + (offs[0] > 0))
        coarse = rng.standard_normal(cshape).astype(np.float32)
        ref = np.zeros(fshape, np.float32)
        O.lib().or_coarse_to_fine(ref, fshape[2], fshape[1], fshape[0], *offs, hf, coarse, cshape[2], cshape[1],
                                  cshape[0], *offs, float(hc))
        got = _t(np.zeros(fshape, np.float32))
        api.coarse_to_fine(got, offs, float(hf), _t(coarse), offs, float(hc))
        assert np.array_equal(bits(_np(got)), bits(ref))
        fine = rng.standard_normal(fshape).astype(np.float32)
        diff = _t(np.zeros(fshape, np.float32))
        api.coarse_to_fine(diff, offs, float(hf), _t(coarse), offs, float(hc), fine_minus=_t(fine))
        assert np.array_equal(bits(_np(diff)), bits(fine - ref))


@pytest.mark.parametrize("cdims,N", CASES)
def test_set_face_boundary_bitexact_and_flux(cuda, cdims, N):
    print(f"set_face_boundary {cdims} N={N}", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    cx, cy, cz = cdims
    rng = np.random.default_rng(20 + N)
    fu = rng.standard_normal((cz * N, cy * N, cx * N + 1)).astype(np.float32)
    cu = rng.standard_normal((cz, cy, cx + 1)).astype(np.float32)
    ref = fu.copy()
    O.lib().or_set_face_boundary(ref, cx * N + 1, cy * N, cz * N, cu, cx + 1, cy, cz, N, 1, N, N)
    got = _t(fu)
    api.set_face_boundary(got, _t(cu), N, (1, N, N))
    g = _np(got)
    assert np.array_equal(bits(g), bits(ref))
    # each coarse face's fine faces now average to the coarse value
    for i in range(cx + 1):
        blk = g[:, :, i * N].reshape(cz, N, cy, N).mean(axis=(1, 3))
        assert np.abs(blk - cu[:, :, i]).max() < 1e-5


def test_bricks_round_trip_and_reference_index(cuda):
    print("bricks", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    for (nx, ny, nz) in ((11, 9, 7), (16, 16, 8), (17, 3, 9), (257, 33, 12)):
        rng = np.random.default_rng(nx)
        a = rng.standard_normal((nz, ny, nx)).astype(np.float32)
        ref = np.zeros(((nx + 7) // 8) * ((ny + 7) // 8) * ((nz + 7) // 8) * 512, np.float32)
        O.lib().or_to_blocked(a, nx, ny, nz, ref)
        b = api.to_bricks(_t(a))
        assert np.array_equal(bits(_np(b)), bits(ref))
        assert np.array_equal(bits(_np(api.from_bricks(b, (nz, ny, nx)))), bits(a))
    # positions written by the reference's own Buffer3D (golden)
    for (nx, ny, nz, i, j, k, idx) in GOLD["blocked_index"]:
        a = np.zeros((nz, ny, nx), np.float32)
        a[k, j, i] = 1.0
        b = _np(api.to_bricks(_t(a)))
        assert b[idx] == 1.0 and b.sum() == 1.0
