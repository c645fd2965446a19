"""Seeded synthetic scenes shared by the tests, smoke(), bench.py and the
learned warm start's training data.

Coordinates follow the reference's internal convention: cell (i,j,k) centre at
(i,j,k)*h, u(i,j,k) at ((i-.5)h, jh, kh), nodal phi(i,j,k) at
((i-.5)h, (j-.5)h, (k-.5)h) (S/Smoke_solver3D.h:299-326, S/Smoke_solver3D.cpp:864).
A config-space position x_cfg (cell centres at (i+.5)h) maps to x_ref = x_cfg - h/2.

Solid SDF convention: phi < 0 inside solids. The plume box has solid walls on
five sides (face weight 0 between the ring cells and the interior) and an open
top (weight 1, Dirichlet p = 0 in the top ring cells).
"""
import numpy as np

MASK32 = 0xFFFFFFFF


def randhash_np(seed):
    """Vectorised randhash (S/util.h:238-246) on uint32 arrays."""
    i = (np.asarray(seed, np.uint64) ^ 0xA3C59AC3) & MASK32
    i = (i * 2654435769) & MASK32
    i ^= i >> 16
    i = (i * 2654435769) & MASK32
    i ^= i >> 16
    i = (i * 2654435769) & MASK32
    return i.astype(np.uint32)


def uniform01(seed, stream, idx):
    """Counter-based U[0,1] draws keyed by (seed, stream, element id)."""
    key = (np.asarray(idx, np.uint64) * 0x9E3779B1 + np.uint64(seed) * 0x85EBCA77 + np.uint64(stream) * 0xC2B2AE3D)
    h = randhash_np(key & MASK32)
    return h.astype(np.float64) / 4294967295.0


def nodal_coords(n, h):
    g = (np.arange(n + 1, dtype=np.float64) - 0.5) * h
    return g


def box_phi(ni, nj, nk, h, open_top=True):
    """Nodal SDF of the five (or six) solid walls one cell inside the domain."""
    x = nodal_coords(ni, h)[None, None, :]
    y = nodal_coords(nj, h)[None, :, None]
    z = nodal_coords(nk, h)[:, None, None]
    phi = np.minimum(x - h, (ni - 2) * h - x)
    phi = np.minimum(phi, y - h)
    if not open_top:
        phi = np.minimum(phi, (nj - 2) * h - y)
    phi = np.minimum(phi, np.minimum(z - h, (nk - 2) * h - z))
    return phi


def sphere_phi(ni, nj, nk, h, center_cfg, r):
    c = np.asarray(center_cfg, np.float64) - 0.5 * h  # config -> reference coordinates
    x = nodal_coords(ni, h)[None, None, :]
    y = nodal_coords(nj, h)[None, :, None]
    z = nodal_coords(nk, h)[:, None, None]
    return np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r


def random_spheres(seed, count=None):
    """1-4 spheres, radius U[0.05,0.15], centres in [0.25,0.75]^3 (config coords)."""
    u = uniform01(seed, 7, np.arange(32))
    if count is None:
        count = 1 + int(u[0] * 4) % 4
    out = []
    for s in range(count):
        c = [0.25 + 0.5 * u[1 + 4 * s + q] for q in range(3)]
        r = 0.05 + 0.1 * u[4 + 4 * s]
        out.append((c, r))
    return out


def solid_phi(n, h, spheres, open_top=True):
    phi = box_phi(n, n, n, h, open_top)
    for c, r in spheres:
        phi = np.minimum(phi, sphere_phi(n, n, n, h, c, r))
    return phi.astype(np.float32)


def mac_zeros(n):
    return (np.zeros((n, n, n + 1), np.float32), np.zeros((n, n + 1, n), np.float32),
            np.zeros((n + 1, n, n), np.float32))


def random_mac(n, seed, scale=1.0):
    """Seeded random MAC velocity in [-scale, scale]."""
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]
    out = []
    for q, shp in enumerate(shapes):
        m = int(np.prod(shp))
        out.append(((uniform01(seed, 100 + q, np.arange(m)) * 2 - 1) * scale).reshape(shp).astype(np.float32))
    return out


def projection_problem(n, seed=0, spheres=None, open_top=True, vel_scale=1.0):
    """Face weights come from the oracle or the GPU; this returns phi, u, v, w, solid velocity."""
    h = 1.0 / n
    if spheres is None:
        spheres = random_spheres(seed)
    phi = solid_phi(n, h, spheres, open_top)
    u, v, w = random_mac(n, seed, vel_scale)
    us, vs, ws = mac_zeros(n)
    return dict(n=n, h=np.float32(h), phi=phi, u=u, v=v, w=w, us=us, vs=vs, ws=ws, spheres=spheres)


def box_problem(ni, nj, nk, seed, h=None, vel_scale=1.0):
    """Projection inputs on an ni x nj x nk box: nodal solid SDF (walls one cell in,
    open top, one sphere at a seeded position), seeded random MAC velocity, zero
    solid velocity."""
    h = 1.0 / max(ni, nj, nk) if h is None else h
    u01 = uniform01(seed, 11, np.arange(4))
    c = (ni * h * (0.3 + 0.4 * u01[0]), nj * h * (0.3 + 0.4 * u01[1]), nk * h * (0.3 + 0.4 * u01[2]))
    phi = np.minimum(box_phi(ni, nj, nk, h), sphere_phi(ni, nj, nk, h, c, (0.1 + 0.1 * u01[3]) * min(ni, nj, nk) * h))
    shapes = [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]
    vel = [((uniform01(seed, 100 + q, np.arange(int(np.prod(shp)))) * 2 - 1) * vel_scale).reshape(shp).astype(np.float32)
           for q, shp in enumerate(shapes)]
    zeros = [np.zeros(shp, np.float32) for shp in shapes]
    return dict(h=np.float32(h), phi=phi.astype(np.float32), u=vel[0], v=vel[1], w=vel[2], us=zeros[0], vs=zeros[1],
                ws=zeros[2])


def block_particles(n, ppc, seed, lo=(0.0, 0.0, 0.0), hi=(0.4, 0.6, 1.0), jitter=1.0):
    """ppc jittered particles in every cell whose centre lies in the config-space
    block [lo, hi] of an n^3 grid (the dam-break fill: [0,.4]x[0,.6]x[0,1]).

    Positions are in reference coordinates (cell (i,j,k) centre at (i,j,k)*h),
    cell-major (x fastest), jitter U[-0.5, 0.5)*jitter*h per axis from the counter
    hash keyed by (seed, axis, particle id). Returns float32 (np, 3).
    """
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    sel = [np.nonzero((c >= lo[a]) & (c <= hi[a]))[0] for a in range(3)]
    k, j, i = np.meshgrid(sel[2], sel[1], sel[0], indexing="ij")
    cells = np.repeat(np.stack([i.ravel(), j.ravel(), k.ravel()], 1), ppc, axis=0)
    idx = np.arange(len(cells))
    jit = np.stack([uniform01(seed, 200 + a, idx) - 0.5 for a in range(3)], 1) * jitter
    return ((cells + jit) * h).astype(np.float32)


def particle_velocities(npart, seed, scale=1.0):
    idx = np.arange(npart)
    return (np.stack([uniform01(seed, 300 + a, idx) * 2 - 1 for a in range(3)], 1) * scale).astype(np.float32)


def ragged_particles(ni, nj, nk, h, seed, max_count=13, bmin=(0.0, 0.0, 0.0), outside=0):
    """Particles with an uneven count per cell (U{0..max_count}), each within 0.45 h of
    its cell centre (i,j,k) h [+ bmin], in a seeded random order, plus `outside`
    particles up to 2 cells beyond the grid. Returns float32 pos (np, 3), vel (np, 3)."""
    ncell = ni * nj * nk
    counts = (uniform01(seed, 400, np.arange(ncell)) * (max_count + 1)).astype(np.int64)
    cell = np.repeat(np.arange(ncell), counts)
    ijk = np.stack([cell % ni, (cell // ni) % nj, cell // (ni * nj)], 1).astype(np.float64)
    idx = np.arange(len(cell))
    jit = np.stack([uniform01(seed, 410 + a, idx) - 0.5 for a in range(3)], 1) * 0.9
    pos = (ijk + jit) * h
    if outside:
        o = np.arange(outside)
        dims = np.array([ni, nj, nk], np.float64)
        far = np.stack([uniform01(seed, 420 + a, o) for a in range(3)], 1) * (dims + 4) - 2
        pos = np.concatenate([pos, far * h])
    pos = (pos + np.asarray(bmin, np.float64)).astype(np.float32)
    perm = np.argsort(uniform01(seed, 430, np.arange(len(pos))), kind="stable")
    pos = np.ascontiguousarray(pos[perm])
    return pos, particle_velocities(len(pos), seed + 1)


def affine_mac(ni, nj, nk, h, coef):
    """MAC grids of the affine velocity comp(x) = coef[comp][0] + coef[comp][1:] . x at the
    face positions (u face (i,j,k) at ((i - 1/2) h, j h, k h), likewise v, w)."""
    out = []
    for comp in range(3):
        shape = [nk, nj, ni]
        shape[2 - comp] += 1
        zz, yy, xx = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        x = [xx * h, yy * h, zz * h]
        x[comp] = x[comp] - 0.5 * h
        c = coef[comp]
        out.append((c[0] + c[1] * x[0] + c[2] * x[1] + c[3] * x[2]).astype(np.float32))
    return out


def nested_flip_case(n=16, ratio=2, box=(0.25, 0.75), seed=0, coarse_vel=None, sub_vel=None, plane=None,
                     solid_vel=(0.0, 0.0, 0.0)):
    """A coarse n^3 grid (h = 1/n) and one subdomain over [box)^3 at `ratio`x (the
    SubDomainSolver members: cells, h, bmin, bmax = bmin + cells h in float), for
    the nested FLIP coupling. Velocities: constant vectors when given, else seeded
    random MAC fields. Solid: the plane phi = x . plane[:3] - plane[3] (unit normal)
    evaluated at the nodes of both grids, else a sphere of radius 0.2 at the centre.
    Solid velocity constant. Returns (coarse, sub) dicts of float32 arrays."""
    h = np.float32(1.0 / n)
    hs = np.float32(h / ratio)
    m = int(round((box[1] - box[0]) * n * ratio))
    bmin = np.array([box[0]] * 3, np.float32)
    bmax = (bmin + np.float32(m) * hs).astype(np.float32)

    def fields(cells, hh, origin, vel, s):
        out = {}
        for comp, name in enumerate("uvw"):
            shape = [cells, cells, cells]
            shape[2 - comp] += 1
            if vel is None:
                out[name] = (np.stack([uniform01(seed + s, 600 + comp, np.arange(int(np.prod(shape))))], 0)[0]
                             .reshape(shape) * 2 - 1).astype(np.float32)
            else:
                out[name] = np.full(shape, vel[comp], np.float32)
            out[name + "s"] = np.full(shape, solid_vel[comp], np.float32)
        g = (np.arange(cells + 1, dtype=np.float64) - 0.5) * float(hh)
        z, y, x = np.meshgrid(g + origin[2], g + origin[1], g + origin[0], indexing="ij")
        if plane is not None:
            phi = x * plane[0] + y * plane[1] + z * plane[2] - plane[3]
        else:
            phi = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) - 0.2
        out["phi"] = phi.astype(np.float32)
        out["us"], out["vs"], out["ws"] = out.pop("us"), out.pop("vs"), out.pop("ws")
        out["h"] = hh
        return out

    coarse = fields(n, h, (0.0, 0.0, 0.0), coarse_vel, 0)
    sub = fields(m, hs, tuple(float(b) for b in bmin), sub_vel, 1)
    sub["bmin"], sub["bmax"] = tuple(float(b) for b in bmin), tuple(float(b) for b in bmax)
    return coarse, sub
