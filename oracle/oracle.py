"""ctypes bindings of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.  Each wrapper names the reference code it
restates; the restatement itself (with file:line citations) is
oracle/vflip_oracle.cpp.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.or_fraction_inside2.restype = c_float
        L.or_fraction_inside2.argtypes = [c_float, c_float]
        L.or_fraction_inside4.restype = c_float
        L.or_fraction_inside4.argtypes = [c_float] * 4
        L.or_face_weights.argtypes = [c_int, c_int, c_int, f32p, f32p, f32p, f32p]
        L.or_assemble.argtypes = [c_int, c_int, c_int, c_float, c_float, c_int] + [f32p] * 9 + [ctypes.c_void_p] + [f64p] * 5
        L.or_liquid_phi.argtypes = [c_int, f32p, c_float, c_int, c_int, c_int, c_float, f32p]
        L.or_reorder.restype = c_int
        L.or_reorder.argtypes = ([c_int, c_int, c_int, c_float, c_int, c_float, c_float, c_float, ctypes.c_uint,
                                 f32p, f32p, i32p, i32p, f32p, f32p, f32p, f32p, c_int, f32p, f32p])
        L.or_reseed_jitter.restype = c_float
        vp = ctypes.c_void_p
        L.or_nested_velocity_at.argtypes = [vp, vp, c_int, f32p, f32p]
        L.or_advect_particles_nested.argtypes = [vp, vp, c_float, c_int, f32p]
        L.or_halo_select.restype = c_int
        L.or_halo_select.argtypes = [vp, c_float, c_int, f32p, f32p, c_int, f32p, f32p]
        L.or_seed_ghost.restype = c_int
        L.or_seed_ghost.argtypes = [vp, vp, ctypes.c_uint, c_int, f32p, f32p]
        L.or_advect_tracers.argtypes = [vp, vp, c_float, c_float, c_int, f32p]
        L.or_emit_sphere.argtypes = [ctypes.c_uint, c_int, c_float, c_float, c_float, c_float, c_float, f32p]
        L.or_emit_sub_tracers.restype = c_int
        L.or_emit_sub_tracers.argtypes = [vp, vp, ctypes.c_uint, c_int, f32p]
        L.or_tracer_frame.argtypes = [c_int, f32p, c_float, f32p]
        L.or_pse_kernel.restype = c_float
        L.or_pse_kernel.argtypes = [f32p, f32p, c_float]
        L.or_pse_diffuse.argtypes = [vp, vp, c_float, c_float, c_int, f32p, f32p, i32p, i32p, f32p]
        L.or_reseed_jitter.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.or_level_dims.restype = c_int
        L.or_level_dims.argtypes = [c_int, c_int, c_int, i32p, c_int]
        L.or_coarsen.argtypes = [c_int, c_int, c_int] + [f64p] * 8
        L.or_matvec.argtypes = [c_int, c_int, c_int] + [f64p] * 6
        L.or_rbgs.argtypes = [c_int, c_int, c_int] + [f64p] * 6 + [c_int]
        L.or_restrict_residual.argtypes = [c_int, c_int, c_int] + [f64p] * 7
        L.or_prolong_add.argtypes = [c_int, c_int, c_int, f64p, f64p]
        L.or_vcycle.argtypes = [c_int, c_int, c_int] + [f64p] * 6 + [c_int] * 3
        L.or_amgpcg.restype = c_int
        L.or_amgpcg.argtypes = ([c_int, c_int, c_int] + [f64p] * 6 + [c_double, c_int,
                                ctypes.POINTER(c_double), ctypes.POINTER(c_int), c_int, c_int, c_int,
                                ctypes.c_void_p])
        L.or_grad_apply.argtypes = [c_int, c_int, c_int, c_float, c_float, c_int, f64p] + [f32p] * 9
        L.or_sample_linear.restype = c_float
        L.or_sample_linear.argtypes = [f32p, c_int, c_int, c_int] + [c_double] * 7
        L.or_sample_cubic.restype = c_float
        L.or_sample_cubic.argtypes = [f32p, c_int, c_int, c_int] + [c_double] * 7
        L.or_sample_cube_lerp.restype = c_float
        L.or_sample_cube_lerp.argtypes = [f32p, c_int, c_int, c_int] + [c_double] * 7 + [f32p]
        L.or_advect_rk2.argtypes = ([c_int, c_int, c_int, c_float, c_float, c_int, f32p, f32p, f32p, f32p,
                                    c_int, c_int, c_int, c_double, c_double, c_double, f32p])
        L.or_hash.restype = c_int
        L.or_hash.argtypes = ([c_int, f32p, c_float, c_int, c_int, c_int, c_int, c_float, c_float, c_float,
                              i32p, i32p, i32p])
        L.or_p2g.argtypes = ([c_int, f32p, f32p, c_int, c_float, c_int, c_int, c_int, i32p, i32p,
                             c_int, c_int, c_int, c_double, c_double, c_double, c_int, c_float, c_float,
                             c_float, f32p, f32p])
        L.or_g2p.argtypes = ([c_int, f32p, f32p, c_float, c_int, c_int, c_int] + [f32p] * 6 +
                             [c_float, c_float, c_float, c_float])
        L.or_fine_to_coarse.argtypes = [f32p, c_int, c_int, c_int, f32p, c_int, c_int, c_int] + [c_int] * 4
        L.or_coarse_to_fine.argtypes = ([f32p, c_int, c_int, c_int, c_double, c_double, c_double, c_float,
                                        f32p, c_int, c_int, c_int, c_double, c_double, c_double, c_double])
        L.or_set_face_boundary.argtypes = [f32p, c_int, c_int, c_int, f32p, c_int, c_int, c_int] + [c_int] * 4
        L.or_to_blocked.argtypes = [f32p, c_int, c_int, c_int, f32p]
        L.or_randhash.restype = ctypes.c_uint
        L.or_randhash.argtypes = [ctypes.c_uint]
        L.or_num_threads.restype = c_int
        L.or_plume_force.argtypes = ([c_int, c_int, c_int] + [c_float] * 6 + [ctypes.c_uint] + [c_float] * 3 +
                                     [f32p] * 3)
        L.or_sample_grad.argtypes = [f32p, c_int, c_int, c_int] + [c_double] * 7 + [f32p]
        L.or_normalize3.argtypes = [f32p]
        L.or_advect_particles.argtypes = ([c_int, f32p, f32p, f32p, f32p, f32p, c_int, c_int, c_int, c_float,
                                          c_float, c_int, c_float, c_float, c_float])
        L.or_get_delta.argtypes = [c_int, f32p, f32p, f32p]
        L.or_inflow_bc.argtypes = [c_int, c_int, c_int, np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS"),
                                   c_float, f32p, f32p, f32p]
        L.or_grid_points.argtypes = [c_int, c_int, c_int, c_float, c_float, c_float, c_float, f32p]
        vp = ctypes.c_void_p
        L.or_mg3d_levels.restype = c_int
        L.or_mg3d_levels.argtypes = [c_int, c_int, c_int, vp]
        L.or_mg3d_rbgs.argtypes = [c_int] * 5 + [vp, vp, c_int, c_double]
        L.or_mg3d_residual.argtypes = [c_int] * 5 + [vp] * 3
        L.or_mg3d_restrict.argtypes = [c_int] * 4 + [vp] * 2
        L.or_mg3d_prolongate.argtypes = [c_int] * 4 + [vp] * 2
        L.or_mg3d_solve.argtypes = [c_int] * 5 + [vp, vp, c_int, c_int]
        _LIB = L
    return _LIB


def ref_lib():
    """oracle/_ref/libvflipref.so (reference sources compiled as-is), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libvflipref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_fraction_inside2.restype = c_float
        R.ref_fraction_inside2.argtypes = [c_float, c_float]
        R.ref_fraction_inside4.restype = c_float
        R.ref_fraction_inside4.argtypes = [c_float] * 4
        R.ref_randhash.restype = ctypes.c_uint
        R.ref_randhash.argtypes = [ctypes.c_uint]
        R.ref_sample.argtypes = ([c_int, f32p, c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                                 c_int, f64p, f32p, f32p, f32p])
        R.ref_blocked_index.restype = ctypes.c_uint
        R.ref_blocked_index.argtypes = [c_int] * 6
        R.ref_normalize.argtypes = [c_int, f32p, f32p]
        R.ref_cell_index.argtypes = [c_int, f32p, c_float, c_int, c_float, c_float, c_float, i32p]
        R.ref_grid_points.argtypes = [c_int, c_int, c_int, c_float, c_double, c_double, c_double, f32p]
        _REF = R
    return _REF


# --------------------------------------------------------------------------- helpers
def c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def face_weights(phi):
    """phi: nodal array shaped (nk+1, nj+1, ni+1). Returns uw, vw, ww (z,y,x arrays)."""
    nk, nj, ni = (s - 1 for s in phi.shape)
    uw = np.zeros((nk, nj, ni + 1), np.float32)
    vw = np.zeros((nk, nj + 1, ni), np.float32)
    ww = np.zeros((nk + 1, nj, ni), np.float32)
    lib().or_face_weights(ni, nj, nk, c32(phi), uw, vw, ww)
    return uw, vw, ww


def assemble(h, u, v, w, uw, vw, ww, us, vs, ws, dt=1.0, sub=False, liquid=None):
    """liquid: optional per-cell float32 mask (1 liquid, 0 air = Dirichlet p = 0)."""
    nk, nj, ni1 = u.shape
    ni = ni1 - 1
    shp = (nk, nj, ni)
    out = [np.zeros(shp, np.float64) for _ in range(5)]
    liq = None if liquid is None else c32(liquid)
    lib().or_assemble(ni, nj, nk, h, dt, int(sub), c32(u), c32(v), c32(w), c32(uw), c32(vw), c32(ww),
                      c32(us), c32(vs), c32(ws), None if liq is None else liq.ctypes.data, *out)
    return out  # diag, ax, ay, az, rhs


def liquid_phi(pos, h, shape, radius):
    """Particle level set at the cell centres of a grid shaped (nk, nj, ni)."""
    nk, nj, ni = shape
    p = c32(pos).reshape(-1)
    out = np.zeros(shape, np.float32)
    lib().or_liquid_phi(p.size // 3, p, h, ni, nj, nk, radius, out)
    return out


def reorder(pos, vel, h, dims, start, order, u, v, w, phi, seed, sub=False, bmin=(0.0, 0.0, 0.0)):
    """reorder_FLIP over or_hash's lists; returns (pos, vel) of the new particle set, (n, 3) each."""
    ni, nj, nk = dims
    pos = c32(pos).reshape(-1)
    vel = c32(vel).reshape(-1)
    if pos.size == 0:
        pos = np.zeros(3, np.float32)
        vel = np.zeros(3, np.float32)
    order = np.ascontiguousarray(order, np.int32)
    if order.size == 0:
        order = np.zeros(1, np.int32)
    cap = int(start[-1]) + 8 * ni * nj * nk
    out_p = np.zeros(3 * max(cap, 1), np.float32)
    out_v = np.zeros(3 * max(cap, 1), np.float32)
    n = lib().or_reorder(ni, nj, nk, h, int(sub), *[float(b) for b in bmin], seed & 0xFFFFFFFF, pos, vel,
                         np.ascontiguousarray(start, np.int32), order, c32(u), c32(v), c32(w), c32(phi), cap,
                         out_p, out_v)
    assert n <= cap
    return out_p[:3 * n].reshape(-1, 3), out_v[:3 * n].reshape(-1, 3)


class _Coarse(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("h", c_float)] + \
        [(f, ctypes.c_void_p) for f in ("u", "v", "w", "phi", "us", "vs", "ws")]


class _Sub(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("h", c_float), ("bmin", c_float * 3),
                ("bmax", c_float * 3)] + [(f, ctypes.c_void_p) for f in ("u", "v", "w", "phi", "us", "vs", "ws")]


def _views(coarse, sub):
    """coarse / sub: dicts with h, u, v, w, phi, optional us, vs, ws (numpy, (z, y, x)); sub also bmin, bmax.
    Returns the ctypes views and the arrays they point into (keep them alive)."""
    keep = []

    def arr(d, k):
        a = d.get(k)
        if a is None:
            return None
        a = c32(a)
        keep.append(a)
        return a.ctypes.data

    nz, ny, nx1 = coarse["u"].shape
    C = _Coarse(nx1 - 1, ny, nz, float(coarse["h"]), *[arr(coarse, k) for k in ("u", "v", "w", "phi", "us", "vs", "ws")])
    snz, sny, snx1 = sub["u"].shape
    S = _Sub(snx1 - 1, sny, snz, float(sub["h"]), (c_float * 3)(*sub["bmin"]), (c_float * 3)(*sub["bmax"]),
             *[arr(sub, k) for k in ("u", "v", "w", "phi", "us", "vs", "ws")])
    return C, S, keep


def nested_velocity(coarse, sub, pos):
    C, S, keep = _views(coarse, sub)
    p = c32(pos).reshape(-1)
    out = np.zeros_like(p)
    lib().or_nested_velocity_at(ctypes.byref(C), ctypes.byref(S), p.size // 3, p, out)
    return out.reshape(-1, 3)


def advect_particles_nested(coarse, sub, dt, pos):
    C, S, keep = _views(coarse, sub)
    out = c32(pos).reshape(-1).copy()
    lib().or_advect_particles_nested(ctypes.byref(C), ctypes.byref(S), dt, out.size // 3, out)
    return out.reshape(-1, 3)


def halo_select(sub, dx, pos, vel):
    C, S, keep = _views(sub, sub)
    p, v = c32(pos).reshape(-1), c32(vel).reshape(-1)
    n0 = p.size // 3
    op, ov = np.zeros(max(3 * n0, 3), np.float32), np.zeros(max(3 * n0, 3), np.float32)
    n = lib().or_halo_select(ctypes.byref(S), dx, n0, p if n0 else np.zeros(3, np.float32),
                             v if n0 else np.zeros(3, np.float32), n0, op, ov)
    return op[:3 * n].reshape(-1, 3), ov[:3 * n].reshape(-1, 3)


def seed_ghost(coarse, sub, seed):
    C, S, keep = _views(coarse, sub)
    cap = 8 * S.nx * S.ny * S.nz
    op, ov = np.zeros(3 * cap, np.float32), np.zeros(3 * cap, np.float32)
    n = lib().or_seed_ghost(ctypes.byref(C), ctypes.byref(S), seed & 0xFFFFFFFF, cap, op, ov)
    return op[:3 * n].reshape(-1, 3), ov[:3 * n].reshape(-1, 3)


def pse_kernel(xi, xj, tau):
    return lib().or_pse_kernel(c32(xi), c32(xj), tau)


def pse_diffuse(coarse, sub, dt, nu, np_real, pos, vel, start, order):
    C, S, keep = _views(coarse, sub)
    p, v = c32(pos).reshape(-1), c32(vel).reshape(-1)
    out = np.zeros(3 * max(np_real, 1), np.float32)
    order = np.ascontiguousarray(order, np.int32)
    lib().or_pse_diffuse(ctypes.byref(C), ctypes.byref(S), dt, nu, np_real, p, v,
                         np.ascontiguousarray(start, np.int32), order if order.size else np.zeros(1, np.int32), out)
    return out[:3 * np_real].reshape(-1, 3)


def advect_tracers(coarse, sub, dt, decay, tracers):
    """advect_tracers of float4 tracers (n, 4); sub None = no subdomain."""
    if sub is None:
        C, _, keep = _views(coarse, dict(coarse, bmin=(0, 0, 0), bmax=(0, 0, 0)))
        Sref = None
    else:
        C, S, keep = _views(coarse, sub)
        Sref = ctypes.byref(S)
    out = c32(tracers).reshape(-1).copy()
    lib().or_advect_tracers(ctypes.byref(C), Sref, dt, decay, out.size // 4, out)
    return out.reshape(-1, 4)


def emit_sphere(seed, n_emit, centre, r, hx):
    out = np.zeros(4 * max(n_emit, 1), np.float32)
    lib().or_emit_sphere(seed & 0xFFFFFFFF, n_emit, *[float(c) for c in centre], r, hx, out)
    return out[:4 * n_emit].reshape(-1, 4)


def emit_sub_tracers(coarse, sub, seed):
    C, S, keep = _views(coarse, sub)
    cap = 4 * S.nx * S.ny * S.nz
    out = np.zeros(4 * cap, np.float32)
    n = lib().or_emit_sub_tracers(ctypes.byref(C), ctypes.byref(S), seed & 0xFFFFFFFF, cap, out)
    return out[:4 * n].reshape(-1, 4)


def tracer_frame(tracers, lx):
    t = c32(tracers).reshape(-1)
    out = np.zeros_like(t)
    lib().or_tracer_frame(t.size // 4, t, lx, out)
    return out.reshape(-1, 4)


def reseed_jitter(seed, key):
    return lib().or_reseed_jitter(seed & 0xFFFFFFFF, key & 0xFFFFFFFF)


def level_dims(ni, nj, nk):
    d = np.zeros(3 * 32, np.int32)
    n = lib().or_level_dims(ni, nj, nk, d, 32)
    return [tuple(int(x) for x in d[3 * l:3 * l + 3]) for l in range(n)]


def coarsen(d, ax, ay, az):
    nk, nj, ni = d.shape
    ci, cj, ck = (int(np.ceil(np.float32(n) / 2.0)) for n in (ni, nj, nk))
    out = [np.zeros((ck, cj, ci), np.float64) for _ in range(4)]
    lib().or_coarsen(ni, nj, nk, c64(d), c64(ax), c64(ay), c64(az), *out)
    return out


def matvec(d, ax, ay, az, x):
    nk, nj, ni = d.shape
    y = np.zeros_like(d)
    lib().or_matvec(ni, nj, nk, c64(d), c64(ax), c64(ay), c64(az), c64(x), y)
    return y


def rbgs(d, ax, ay, az, b, x, sweeps):
    nk, nj, ni = d.shape
    x = c64(x).copy()
    lib().or_rbgs(ni, nj, nk, c64(d), c64(ax), c64(ay), c64(az), c64(b), x, sweeps)
    return x


def vcycle(d, ax, ay, az, r, pre=4, post=4, coarse=200):
    nk, nj, ni = d.shape
    z = np.zeros_like(d)
    lib().or_vcycle(ni, nj, nk, c64(d), c64(ax), c64(ay), c64(az), c64(r), z, pre, post, coarse)
    return z


def amgpcg(d, ax, ay, az, rhs, tol=1e-6, maxit=50, pre=4, post=4, coarse=200, history=False):
    nk, nj, ni = d.shape
    x = np.zeros_like(d)
    res = c_double(0)
    its = c_int(0)
    hist = np.zeros(maxit + 1, np.float64) if history else None
    ok = lib().or_amgpcg(ni, nj, nk, c64(d), c64(ax), c64(ay), c64(az), c64(rhs), x, tol, maxit,
                         ctypes.byref(res), ctypes.byref(its), pre, post, coarse,
                         hist.ctypes.data if history else None)
    out = dict(x=x, ok=bool(ok), residual=res.value, iterations=its.value)
    if history:
        out["history"] = hist[: its.value + 1]
    return out


def grad_apply(h, p, uw, vw, ww, us, vs, ws, u, v, w, dt=1.0, sub=False):
    nk, nj, ni = p.shape
    u, v, w = c32(u).copy(), c32(v).copy(), c32(w).copy()
    lib().or_grad_apply(ni, nj, nk, h, dt, int(sub), c64(p), c32(uw), c32(vw), c32(ww), c32(us), c32(vs),
                        c32(ws), u, v, w)
    return u, v, w


def inflow_bc(bdesc, u_in, u, v, w):
    """S/Smoke_solver3D.cpp:964-980 on copies of u, v, w; bdesc int8 (nk, nj, ni)."""
    nk, nj, ni = bdesc.shape
    u, v, w = c32(u).copy(), c32(v).copy(), c32(w).copy()
    lib().or_inflow_bc(ni, nj, nk, np.ascontiguousarray(bdesc, np.int8), u_in, u, v, w)
    return u, v, w


def grid_points(shape, h, offs):
    """World positions (float32, shape + (3,)) where set_boundary_phi/vel sample their
    callbacks on a grid shaped (nz, ny, nx) with grid offsets offs = (ox, oy, oz)."""
    nz, ny, nx = shape
    out = np.zeros((nz, ny, nx, 3), np.float32)
    lib().or_grid_points(nx, ny, nz, h, offs[0], offs[1], offs[2], out)
    return out


def advect_rk2(h, dt, u, v, w, field, offs, mode=0):
    nk, nj, ni1 = u.shape
    ni = ni1 - 1
    fz, fy, fx = field.shape
    out = np.zeros_like(c32(field))
    lib().or_advect_rk2(ni, nj, nk, h, dt, mode, c32(u), c32(v), c32(w), c32(field), fx, fy, fz,
                        offs[0], offs[1], offs[2], out)
    return out


def bin_particles(pos, h, dims, clamp=False, bmin=(0.0, 0.0, 0.0)):
    ni, nj, nk = dims
    pos = c32(pos).reshape(-1)
    npart = pos.size // 3
    key = np.zeros(npart, np.int32)
    start = np.zeros(ni * nj * nk + 1, np.int32)
    order = np.zeros(max(npart, 1), np.int32)
    nb = lib().or_hash(npart, pos, h, ni, nj, nk, int(clamp), *[float(b) for b in bmin], key, start, order)
    return key, start, order[:nb]


def p2g(pos, vel, comp, h, dims, start, order, fshape, offs, sub=False, bmin=(0.0, 0.0, 0.0)):
    ni, nj, nk = dims
    fz, fy, fx = fshape
    field = np.zeros(fshape, np.float32)
    coef = np.zeros(fshape, np.float32)
    pos = c32(pos).reshape(-1)
    order = np.ascontiguousarray(order, np.int32)
    if order.size == 0:
        order = np.zeros(1, np.int32)
    lib().or_p2g(pos.size // 3, pos, c32(vel).reshape(-1), comp, h, ni, nj, nk,
                 np.ascontiguousarray(start, np.int32), order, fx, fy, fz,
                 offs[0], offs[1], offs[2], int(sub), *[float(b) for b in bmin], field, coef)
    return field, coef


def g2p(pos, vel, h, dims, du, dv, dw, u=None, v=None, w=None, flip=1.0, bmin=(0.0, 0.0, 0.0)):
    ni, nj, nk = dims
    vel = c32(vel).reshape(-1).copy()
    pos = c32(pos).reshape(-1)
    u = du if u is None else u
    v = dv if v is None else v
    w = dw if w is None else w
    lib().or_g2p(pos.size // 3, pos, vel, h, ni, nj, nk, c32(du), c32(dv), c32(dw), c32(u), c32(v), c32(w),
                 flip, *[float(b) for b in bmin])
    return vel.reshape(-1, 3)


def randhash(s):
    return lib().or_randhash(s)


def plume_force(h, dt, center, r, seed, jitter, alpha, beta, rho, temp, v):
    nk, nj, ni = rho.shape
    rho, temp, v = c32(rho).copy(), c32(temp).copy(), c32(v).copy()
    lib().or_plume_force(ni, nj, nk, h, dt, center[0], center[1], center[2], r, seed, jitter, alpha, beta, rho, temp,
                         v)
    return rho, temp, v


def sample_grad(f, h, offs, pts):
    """Buffer3D::sample_grad at each world point of pts (n,3 float64)."""
    fz, fy, fx = f.shape
    f = c32(f)
    out = np.zeros((len(pts), 3), np.float32)
    g = np.zeros(3, np.float32)
    for n, (x, y, z) in enumerate(np.asarray(pts, np.float64)):
        lib().or_sample_grad(f, fx, fy, fz, h, offs[0], offs[1], offs[2], x, y, z, g)
        out[n] = g
    return out


def normalize(vecs):
    out = c32(vecs).reshape(-1, 3).copy()
    for n in range(len(out)):
        row = np.ascontiguousarray(out[n])
        lib().or_normalize3(row)
        out[n] = row
    return out


def advect_particles(pos, h, dt, u, v, w, phi, sub=False, bmin=(0.0, 0.0, 0.0)):
    """traceRK3 + solid push-out (advect_FLIP) of xyz positions; returns new positions."""
    nk, nj, ni1 = u.shape
    out = c32(pos).reshape(-1).copy()
    lib().or_advect_particles(out.size // 3, out, c32(u), c32(v), c32(w), c32(phi), ni1 - 1, nj, nk, h, dt,
                              int(sub), *[float(b) for b in bmin])
    return out.reshape(-1, 3)


def get_delta(new, old):
    out = np.zeros_like(c32(new))
    lib().or_get_delta(out.size, c32(new), c32(old), out)
    return out


# --------------------------------------------------------------------------- geometric multigrid
# MultiGridSolver3D<T> (S/Multigrid3D.h), arrays (nz, ny, nx) of float32 (MGSolverf) or
# float64 (MGSolverd); the restatement is oracle/vflip_oracle.cpp, namespace mg3d.
def _mg_type(a):
    if a.dtype == np.float32:
        return 0
    if a.dtype == np.float64:
        return 1
    raise ValueError("float32 or float64 arrays")


def _vp(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def mg3d_levels(nx, ny, nz):
    """[(nx, ny, nz)] per level (MultiGrid::m_ComputeLevels / m_AssignMemoryEachLevel)."""
    n = lib().or_mg3d_levels(nx, ny, nz, None)
    d = np.zeros(3 * n, np.int32)
    lib().or_mg3d_levels(nx, ny, nz, _vp(d))
    return [tuple(int(v) for v in d[3 * l:3 * l + 3]) for l in range(n)]


def mg3d_rbgs(b, x, iters, omega=1.0, dirichlet=False):
    """m_Red_Black_Gauss: returns the smoothed copy of x."""
    x = np.ascontiguousarray(x).copy()
    b = np.ascontiguousarray(b, dtype=x.dtype)
    nz, ny, nx = x.shape
    lib().or_mg3d_rbgs(nx, ny, nz, _mg_type(x), int(dirichlet), _vp(b), _vp(x), int(iters), float(omega))
    return x


def mg3d_residual(b, x, dirichlet=False):
    x = np.ascontiguousarray(x)
    b = np.ascontiguousarray(b, dtype=x.dtype)
    res = np.empty_like(x)
    nz, ny, nx = x.shape
    lib().or_mg3d_residual(nx, ny, nz, _mg_type(x), int(dirichlet), _vp(res), _vp(b), _vp(x))
    return res


def mg3d_restrict(src):
    src = np.ascontiguousarray(src)
    nz, ny, nx = src.shape
    dst = np.empty((max(nz // 2, 1), max(ny // 2, 1), max(nx // 2, 1)), src.dtype)
    lib().or_mg3d_restrict(nx, ny, nz, _mg_type(src), _vp(src), _vp(dst))
    return dst


def mg3d_prolongate(src, dst):
    """m_Prolongate: returns dst + the parents' values (dst is not modified)."""
    dst = np.ascontiguousarray(dst).copy()
    src = np.ascontiguousarray(src, dtype=dst.dtype)
    nz, ny, nx = dst.shape
    assert src.shape == (max(nz // 2, 1), max(ny // 2, 1), max(nx // 2, 1))
    lib().or_mg3d_prolongate(nx, ny, nz, _mg_type(dst), _vp(src), _vp(dst))
    return dst


def mg3d_solve(b, x, cycles=1, fmg=False, dirichlet=False):
    """One solver object: m_FullMultiGrid first when fmg, then `cycles` m_Vcycle calls;
    returns the new x."""
    x = np.ascontiguousarray(x).copy()
    b = np.ascontiguousarray(b, dtype=x.dtype)
    nz, ny, nx = x.shape
    lib().or_mg3d_solve(nx, ny, nz, _mg_type(x), int(dirichlet), _vp(x), _vp(b), int(cycles), int(bool(fmg)))
    return x
