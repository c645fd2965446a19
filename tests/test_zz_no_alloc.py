"""vflip_hash_noalloc (the binning with its scratch in a static device array, no
allocation) against vflip_hash (the same kernels over a workspace object), through
the C ABI: bit-identical keys, cell offsets and per-cell order on two calls in a
row, the second with fewer particles on a smaller grid (the array keeps the first
call's contents, so anything not cleared would show), plus the size limits, which
are refused before any launch.
"""
import ctypes

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def _run_noalloc(L, dims, h, pos, clamp=0, bmin=(0.0, 0.0, 0.0)):
    import torch
    ni, nj, nk = dims
    npart = pos.shape[0]
    key = torch.full((max(npart, 1),), -7, dtype=torch.int32, device="cuda")
    start = torch.full((ni * nj * nk + 1,), -7, dtype=torch.int32, device="cuda")
    order = torch.full((max(npart, 1),), -7, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.vflip_hash_noalloc(ni, nj, nk, h, npart, ctypes.c_void_p(pos.data_ptr()), clamp, *bmin,
                              ctypes.c_void_p(key.data_ptr()), ctypes.c_void_p(start.data_ptr()),
                              ctypes.c_void_p(order.data_ptr()), stream)
    assert rc == 0, L.vflip_last_error()
    torch.cuda.synchronize()
    return key[:npart].cpu().numpy(), start.cpu().numpy(), order[:npart].cpu().numpy()


def test_hash_noalloc_matches_hash(cuda):
    import torch
    from vflip_amd import api
    from vflip_amd._lib import lib
    L = lib()
    cases = [((48, 40, 36), 11, 9), ((20, 18, 16), 4, 3)]  # second: smaller grid, fewer particles
    for (dims, seed, max_count) in cases:
        ni, nj, nk = dims
        h = np.float32(1.0 / 48)
        print(f"no-alloc binning {dims} seed {seed}", flush=True)
        pos_np, _ = scenes.ragged_particles(ni, nj, nk, h, seed, max_count=max_count, outside=200)
        # a pile-up: 300 particles in one cell (the long-segment sort)
        pile = np.repeat(pos_np[:1], 300, axis=0)
        pos_np = np.concatenate([pos_np, pile]).astype(np.float32)
        pos = torch.from_numpy(pos_np).cuda()
        for clamp in (0, 1):
            P = api.Particles(ni, nj, nk, float(h))
            k_ref, s_ref, o_ref, _ = P.hash(pos, clamp=bool(clamp))
            k, s, o = _run_noalloc(L, dims, float(h), pos, clamp)
            assert np.array_equal(k, k_ref.cpu().numpy())
            assert np.array_equal(s, s_ref.cpu().numpy())
            nb = int(s[-1])  # binned particles (dropped ones have no order entry)
            assert np.array_equal(o[:nb], o_ref[:nb].cpu().numpy())
            del P
    # limits: refused before any launch
    rc = L.vflip_hash_noalloc(8, 8, 8, 0.1, (1 << 22) + 1, None, 0, 0.0, 0.0, 0.0, None, None, None, None)
    assert rc == -1 and b"no-alloc limit" in L.vflip_last_error()
    rc = L.vflip_hash_noalloc(512, 256, 129, 0.1, 0, None, 0, 0.0, 0.0, 0.0, None, None, None, None)
    assert rc == -1 and b"no-alloc limit" in L.vflip_last_error()
