"""GPU FLIP transfer against the oracle.

Bit-exact: binning (keys, cell offsets, per-cell order), the exact-order P2G
gather, getDelta, G2P and particle advection. The LDS-binned scatter P2G sums
the same float weights in another order, so it is held to FP32 rounding
(relative L2 <= 1e-6 against the oracle, tolerance stated per test).
"""
import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _cloud(n, ppc, seed, spread=1.0, shuffle=False, bmin=(0.0, 0.0, 0.0)):
    """ppc particles per cell over the whole grid; spread > 1 pushes some out of the grid."""
    pos = scenes.block_particles(n, ppc, seed, hi=(1.0, 1.0, 1.0), jitter=spread)
    pos = pos + np.asarray(bmin, np.float32)
    vel = scenes.particle_velocities(len(pos), seed)
    if shuffle:
        perm = np.random.default_rng(seed).permutation(len(pos))
        pos, vel = pos[perm], vel[perm]
    return np.ascontiguousarray(pos), np.ascontiguousarray(vel)


def _hash_both(n, pos, clamp=False, bmin=(0.0, 0.0, 0.0), h=None):
    from oracle import oracle as O
    from vflip_amd import api
    h = np.float32(1.0 / n) if h is None else h
    P = api.Particles(n, n, n, float(h))
    key, start, order, nb = P.hash(_t(pos), clamp=clamp, bmin=bmin)
    rk, rs, ro = O.bin_particles(pos, h, (n, n, n), clamp=clamp, bmin=bmin)
    return P, (_np(key), _np(start), _np(order)[:nb], nb), (rk, rs, ro)


@pytest.mark.parametrize("n,clamp,shuffle", [(20, False, False), (20, False, True), (33, True, True),
                                             (33, True, False)])
def test_hash_bitexact(cuda, n, clamp, shuffle):
    print(f"hash n={n} clamp={clamp} shuffle={shuffle}", flush=True)
    bmin = (0.1, 0.05, 0.2) if clamp else (0.0, 0.0, 0.0)
    pos, _ = _cloud(n, 8, n, spread=1.8, shuffle=shuffle, bmin=bmin)
    _, (k, s, o, nb), (rk, rs, ro) = _hash_both(n, pos, clamp, bmin)
    assert np.array_equal(k, rk)
    assert np.array_equal(s, rs)
    assert nb == len(ro) and np.array_equal(o, ro)
    if not clamp:
        assert nb < len(pos)  # some particles were dropped


def test_hash_long_segments(cuda):
    """Clamping piles thousands of far-away particles into a few boundary cells:
    segments longer than the per-thread fix-up take the workgroup merge sort."""
    print("hash long segments", flush=True)
    n = 16
    rng = np.random.default_rng(3)
    pos, _ = _cloud(n, 4, 1, shuffle=True)
    far = np.concatenate([rng.uniform(-5, -1, (3000, 3)), rng.uniform(2, 3, (1500, 3)),
                          np.stack([rng.uniform(0, 1, 5000), np.full(5000, -2.0), np.full(5000, 9.0)], 1)])
    pos = np.ascontiguousarray(np.concatenate([pos, far.astype(np.float32)])[rng.permutation(len(pos) + len(far))])
    _, (k, s, o, nb), (rk, rs, ro) = _hash_both(n, pos, clamp=True)
    assert np.diff(rs).max() > 1000
    assert np.array_equal(k, rk) and np.array_equal(s, rs) and np.array_equal(o, ro)


def test_hash_empty_and_all_dropped(cuda):
    print("hash empty", flush=True)
    from vflip_amd import api
    n = 8
    P = api.Particles(n, n, n, 1.0 / n)
    pos = _t(np.zeros((0, 3), np.float32))
    key, start, order, nb = P.hash(pos)
    assert nb == 0 and not _np(start).any()
    out = np.full((100, 3), 5.0, np.float32)
    out[:50] = -3.0
    out[0] = np.nan
    key, start, order, nb = P.hash(_t(out))
    assert nb == 0 and np.all(_np(key) == -1) and not _np(start).any()


def test_hash_dam_break_128(cuda):
    """The 128^3 dam-break fill (4.02M particles), shuffled: full-size keys,
    offsets and order against the oracle, and order is a stable sort by key."""
    print("hash dam break 128", flush=True)
    n = 128
    pos = scenes.block_particles(n, 8, 0)
    perm = np.random.default_rng(0).permutation(len(pos))
    pos = np.ascontiguousarray(pos[perm])
    _, (k, s, o, nb), (rk, rs, ro) = _hash_both(n, pos)
    assert nb == len(pos) == 4021248
    assert np.array_equal(k, rk) and np.array_equal(s, rs) and np.array_equal(o, ro)
    assert np.array_equal(o, np.argsort(k, kind="stable"))


def _p2g_case(variant):
    if variant == 0:
        n, bmin = 20, (0.0, 0.0, 0.0)
        pos, vel = _cloud(n, 8, 11, spread=1.4, shuffle=True)
    else:
        n, bmin = 24, (0.3, 0.25, 0.1)
        pos, vel = _cloud(n, 8, 12, spread=1.2, shuffle=True, bmin=bmin)
    return n, bmin, pos, vel


def _oracle_p2g(n, pos, vel, start, order, sub, bmin):
    from oracle import oracle as O
    h = np.float32(1.0 / n)
    out = []
    for comp, fshape, offs in ((0, (n, n, n + 1), (0.5, 0, 0)), (1, (n, n + 1, n), (0, 0.5, 0)),
                               (2, (n + 1, n, n), (0, 0, 0.5))):
        out.append(O.p2g(pos, vel, comp, h, (n, n, n), start, order, fshape, offs, sub=sub, bmin=bmin))
    return out


@pytest.mark.parametrize("variant", [0, 1])
def test_p2g_exact_order_bitexact(cuda, variant):
    print(f"p2g exact variant={variant}", flush=True)
    from vflip_amd import api
    n, bmin, pos, vel = _p2g_case(variant)
    P, (k, s, o, nb), (rk, rs, ro) = _hash_both(n, pos, clamp=variant == 1, bmin=bmin)
    gs, go = _t(s), _t(np.concatenate([o, np.zeros(1, np.int32)]))
    got = P.p2g(_t(pos), _t(vel), gs, go, variant=variant, method=1, bmin=bmin)
    ref = _oracle_p2g(n, pos, vel, rs, ro, variant == 1, bmin)
    for (gf, gc), (rf, rc) in zip(got, ref):
        assert np.array_equal(bits(_np(gc)), bits(rc))
        assert np.array_equal(bits(_np(gf)), bits(rf))


@pytest.mark.parametrize("variant,method", [(0, 0), (1, 0), (0, 2), (1, 2)])
def test_p2g_scatter_matches_oracle(cuda, variant, method):
    """Owner-computes (method 0) and global-atomic (method 2) scatters differ from
    the reference at FP32 rounding level. Method 2 sums the reference's own
    weights in another order: relative L2 <= 1e-6. Method 0 also rounds each
    weight differently (one scaled offset per axis, about 6e-8 absolute per
    weight, as large as the reference's own rounding): relative L2 <= 1e-5, the
    FP32 field tolerance. Both: max |diff| <= 1e-5 * max |ref|."""
    print(f"p2g scatter variant={variant} method={method}", flush=True)
    n, bmin, pos, vel = _p2g_case(variant)
    P, (k, s, o, nb), (rk, rs, ro) = _hash_both(n, pos, clamp=variant == 1, bmin=bmin)
    gs, go = _t(s), _t(np.concatenate([o, np.zeros(1, np.int32)]))
    got = P.p2g(_t(pos), _t(vel), gs, go, variant=variant, method=method, bmin=bmin)
    ref = _oracle_p2g(n, pos, vel, rs, ro, variant == 1, bmin)
    tol = 1e-5 if method == 0 else 1e-6
    for (gf, gc), (rf, rc) in zip(got, ref):
        for g, r in ((_np(gc), rc), (_np(gf), rf)):
            assert np.linalg.norm(g - r) <= tol * np.linalg.norm(r)
            assert np.abs(g - r).max() <= 1e-5 * np.abs(r).max()


def test_p2g_scatter_dam_break_128(cuda):
    """Full-size scatter P2G (4.02M particles, 128^3, both scatter methods)
    against the exact-order gather; relative L2 <= 1e-5 (method 0) / 1e-6 (method 2)."""
    print("p2g scatter 128", flush=True)
    from vflip_amd import api
    n = 128
    pos = scenes.block_particles(n, 8, 0)
    vel = scenes.particle_velocities(len(pos), 0)
    P = api.Particles(n, n, n, 1.0 / n)
    gp, gv = _t(pos), _t(vel)
    key, start, order, nb = P.hash(gp)
    exact = P.p2g(gp, gv, start, order, method=1)
    for method in (0, 2):
        fast = P.p2g(gp, gv, start, order, method=method)
        for (ff, fc), (ef, ec) in zip(fast, exact):
            for g, r in ((_np(fc), _np(ec)), (_np(ff), _np(ef))):
                assert np.linalg.norm(g - r) <= (1e-5 if method == 0 else 1e-6) * np.linalg.norm(r)


def test_get_delta_and_g2p_bitexact(cuda):
    print("g2p", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    n = 24
    h = np.float32(1.0 / n)
    new = scenes.random_mac(n, 5)
    old = scenes.random_mac(n, 6)
    pos, vel = _cloud(n, 4, 7, spread=1.5, shuffle=True)
    d = [O.get_delta(a, b) for a, b in zip(new, old)]
    gd = [api.get_delta(_t(a), _t(b)) for a, b in zip(new, old)]
    for g, r in zip(gd, d):
        assert np.array_equal(bits(_np(g)), bits(r))
    for flip, with_old in ((1.0, True), (1.0, False), (0.95, True)):
        gv = _t(vel)
        if with_old:
            api.g2p(float(h), _t(pos), gv, [_t(a) for a in new], [_t(a) for a in old], flip=flip)
        else:
            api.g2p(float(h), _t(pos), gv, gd, None, flip=flip)
        ref = O.g2p(pos, vel, h, (n, n, n), *d, u=new[0], v=new[1], w=new[2], flip=flip)
        assert np.array_equal(bits(_np(gv)), bits(ref)), (flip, with_old)
    # subdomain form: positions relative to bmin
    gv = _t(vel)
    api.g2p(float(h), _t(pos), gv, [_t(a) for a in new], [_t(a) for a in old], bmin=(0.1, 0.2, 0.05))
    ref = O.g2p(pos, vel, h, (n, n, n), *d, bmin=(0.1, 0.2, 0.05))
    assert np.array_equal(bits(_np(gv)), bits(ref))


@pytest.mark.parametrize("variant", [0, 1])
def test_advect_particles_bitexact(cuda, variant):
    print(f"advect particles variant={variant}", flush=True)
    from oracle import oracle as O
    from vflip_amd import api
    n = 32
    h = np.float32(1.0 / n)
    u, v, w = scenes.random_mac(n, 9, scale=2.0)
    phi = scenes.solid_phi(n, 1.0 / n, [((0.5, 0.5, 0.5), 0.2)])
    bmin = (0.05, 0.1, 0.02) if variant == 1 else (0.0, 0.0, 0.0)
    pos, _ = _cloud(n, 2, 13, spread=1.0, shuffle=True)
    dt = np.float32(0.02)
    gp = _t(pos)
    api.advect_particles(float(h), float(dt), gp, _t(u), _t(v), _t(w), _t(phi), variant=variant, bmin=bmin)
    ref = O.advect_particles(pos, h, dt, u, v, w, phi, sub=variant == 1, bmin=bmin)
    got = _np(gp)
    assert np.array_equal(bits(got), bits(ref))
    # the solid push-out changed a good number of particles
    free = O.advect_particles(pos, h, dt, u, v, w, np.ones_like(phi), sub=variant == 1, bmin=bmin)
    assert (ref != free).any(axis=1).sum() > 20


# ------------------------------------------------------------------ reorder_FLIP
def _reorder_both(dims, h, pos, vel, grids, phi, seed, sub=False, bmin=(0.0, 0.0, 0.0), out=None):
    from oracle import oracle as O
    from vflip_amd import api
    P = api.Particles(*dims, float(h))
    gp, gv = _t(pos), _t(vel)
    _, start, order, _ = P.hash(gp, clamp=sub, bmin=bmin)
    op, ov = P.reorder(gp, gv, start, order, *[_t(g) for g in grids], _t(phi), seed,
                       variant=api.SUB if sub else api.COARSE, bmin=bmin, out=out)
    _, rs, ro = O.bin_particles(pos, h, dims, clamp=sub, bmin=bmin)
    rp, rv = O.reorder(pos, vel, h, dims, rs, ro, *grids, phi, seed, sub=sub, bmin=bmin)
    return P, (_np(op), _np(ov)), (rp, rv)


@pytest.mark.parametrize("sub", [False, True])
def test_reorder_bitexact(cuda, sub):
    """Ragged cells (0..13 particles, some outside the grid), random MAC velocity, a solid
    sphere: the GPU rebuild equals the oracle's bit for bit (order, kept and new particles)."""
    print(f"reorder sub={sub}", flush=True)
    dims, h = (20, 18, 16), np.float32(0.05)
    bmin = (0.3, -0.2, 0.1) if sub else (0.0, 0.0, 0.0)
    pos, vel = scenes.ragged_particles(*dims, float(h), seed=6, bmin=bmin, outside=200)
    rng = np.random.default_rng(3)
    grids = [rng.standard_normal(s).astype(np.float32) for s in
             [(dims[2], dims[1], dims[0] + 1), (dims[2], dims[1] + 1, dims[0]), (dims[2] + 1, dims[1], dims[0])]]
    phi = scenes.sphere_phi(*dims, float(h), (0.5, 0.45, 0.4), 0.22).astype(np.float32)
    _, (gp, gv), (rp, rv) = _reorder_both(dims, h, pos, vel, grids, phi, 11, sub=sub, bmin=bmin)
    assert len(gp) == len(rp) and len(rp) > 8 * 1000
    assert np.array_equal(bits(gp), bits(rp)) and np.array_equal(bits(gv), bits(rv))
    free = _reorder_both(dims, h, pos, vel, grids, np.ones_like(phi), 11, sub=sub, bmin=bmin)[2][0]
    assert len(free) - len(rp) > 500  # the sphere dropped particles


def test_reorder_capacity_and_empty(cuda):
    """An output buffer too small is reported and replaced; zero particles give 8 new
    particles in every cell outside the solid."""
    from vflip_amd import api
    dims, h = (12, 10, 9), np.float32(1.0 / 12)
    pos, vel = scenes.ragged_particles(*dims, float(h), seed=8)
    grids = scenes.affine_mac(*dims, float(h), [(0.1, 1, 0, 0), (0.2, 0, 1, 0), (0.3, 0, 0, 1)])
    phi = scenes.box_phi(*dims, float(h), open_top=False).astype(np.float32)
    import torch
    small = (torch.empty((10, 3), device="cuda"), torch.empty((10, 3), device="cuda"))
    _, (gp, gv), (rp, rv) = _reorder_both(dims, h, pos, vel, grids, phi, 4, out=small)
    assert len(gp) > 10 and np.array_equal(bits(gp), bits(rp)) and np.array_equal(bits(gv), bits(rv))
    z = np.zeros((0, 3), np.float32)
    P, (gp, gv), (rp, rv) = _reorder_both(dims, h, z, z, grids, phi, 4)
    assert np.array_equal(bits(gp), bits(rp)) and np.array_equal(bits(gv), bits(rv))
    inside = (phi[:-1, :-1, :-1] > 0)  # cells well inside the walls keep all 8
    assert len(gp) >= 8 * int(inside[2:-2, 2:-2, 2:-2].sum())
    with pytest.raises(api.VflipError):
        from vflip_amd._lib import lib
        import ctypes
        n = ctypes.c_int(0)
        zc = torch.zeros(P.ncell + 1, dtype=torch.int32, device="cuda")
        g = [_t(a) for a in grids]
        api.check(lib().vflip_reorder(P.hnd, 7, 0.0, 0.0, 0.0, 0, 0, None, None, api._ptr(zc, torch.int32), None,
                                      *[api._ptr(a) for a in g], api._ptr(_t(phi)), 0, None, None,
                                      ctypes.byref(n)), "reorder")


def test_reorder_smoke_128(cuda):
    """128^3 FLIP smoke set (8 jittered particles per cell spread over 1.6 cells, shuffled),
    closed box with a sphere: bit-exact with the oracle at full size; the result is cell
    sorted with 8 or 9 particles in every cell away from the solid, and a second rebuild
    without the solid changes nothing."""
    print("reorder 128^3", flush=True)
    from vflip_amd import api
    n = 128
    h = np.float32(1.0 / n)
    pos, vel = _cloud(n, 8, 21, spread=1.6, shuffle=True)
    grids = scenes.random_mac(n, 5, scale=1.0)
    phi = scenes.solid_phi(n, 1.0 / n, [((0.5, 0.5, 0.5), 0.15)], open_top=False)
    P, (gp, gv), (rp, rv) = _reorder_both((n, n, n), h, pos, vel, grids, phi, 17)
    assert np.array_equal(bits(gp), bits(rp)) and np.array_equal(bits(gv), bits(rv))
    tp, tv = _t(gp), _t(gv)
    _, start, order, nb = P.hash(tp)
    counts = _np(start)[1:] - _np(start)[:-1]
    assert nb == len(gp) and np.array_equal(_np(order)[:nb], np.arange(nb))
    cphi = phi[:-1, :-1, :-1].reshape(-1)  # nodal phi at the cell's lower corner
    away = cphi > 2.0 * float(h)
    assert counts[away].min() >= 8 and counts[away].max() <= 9
    ones = _t(np.ones_like(phi))
    g = [_t(a) for a in grids]
    p2, v2 = P.reorder(tp, tv, start, order, *g, ones, 99)
    p2, v2 = _np(p2), _np(v2)
    keep = (counts >= 8)  # no cell rebuilt except those the solid thinned
    assert len(p2) == len(gp) + int((8 - counts[~keep]).sum())


def test_reorder_face_particles_sparse_solid(cuda):
    """Particles within a few ulp of the cell faces (h = 0.1, so the float binning and the
    double sampler round differently) and a solid of isolated negative nodes: the GPU's
    solid-free cell test never changes a decision (bit-exact with the oracle, which samples
    every candidate)."""
    print("reorder face particles", flush=True)
    dims, h = (14, 12, 10), np.float32(0.1)
    rng = np.random.default_rng(12)
    ncell = dims[0] * dims[1] * dims[2]
    cell = rng.integers(0, ncell, 6000)
    ijk = np.stack([cell % dims[0], (cell // dims[0]) % dims[1], cell // (dims[0] * dims[1])], 1)
    face = (ijk + rng.choice([-0.5, 0.5], ijk.shape)) * h
    pos = np.nextafter(face.astype(np.float32), np.float32(np.inf) * rng.choice([-1, 1], ijk.shape))
    pos = np.where(rng.random(ijk.shape) < 0.5, face.astype(np.float32), pos).astype(np.float32)
    pos = np.ascontiguousarray(pos)
    vel = scenes.particle_velocities(len(pos), 3)
    # tiny positive nodes and isolated very negative ones: a face particle whose sampler base
    # node is one below its cell is dropped through the far node's tiny weight (94 of the
    # 382 dropped particles here would survive a test of the cell's own 2x2x2 nodes only)
    phi = np.full((dims[2] + 1, dims[1] + 1, dims[0] + 1), 1e-6, np.float32)
    phi.reshape(-1)[rng.integers(0, phi.size, 40)] = -1e3
    grids = scenes.affine_mac(*dims, float(h), [(0.1, 1, 0, 0), (0.2, 0, 1, 0), (0.3, 0, 0, 1)])
    _, (gp, gv), (rp, rv) = _reorder_both(dims, h, pos, vel, grids, phi, 5)
    assert np.array_equal(bits(gp), bits(rp)) and np.array_equal(bits(gv), bits(rv))
    free = _reorder_both(dims, h, pos, vel, grids, np.ones_like(phi), 5)[2][0]
    assert len(free) > len(rp)
