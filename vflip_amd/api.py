"""Thin Python view of the C ABI over torch device tensors.

torch only provides device memory and the current stream here; every
computation runs in libvflip.so's HIP kernels. Arrays are torch tensors shaped
(nk, nj, ni) (z, y, x) so that the memory layout is the ABI's linear x-fastest.
"""
import ctypes
import os

import torch

from ._lib import (ERR_CAPACITY, PHI_FN, VEL_FN, CoarseView, SolveOpts, SolveStats, SubdomainView, VflipError,  # noqa: F401
                   check, lib)

FP32 = 0
MIXED = 1
COARSE = 0
SUB = 1


def _ptr(t, dtype=torch.float32, numel=None, name="array"):
    if t is None:
        return None
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    return ctypes.c_void_p(t.data_ptr())


def _mac_ptrs(ni, nj, nk, arrays, names):
    """Pointers of MAC-staggered arrays (u, v, w order repeated), each checked against its
    face count: u (ni+1) nj nk, v ni (nj+1) nk, w ni nj (nk+1)."""
    counts = ((ni + 1) * nj * nk, ni * (nj + 1) * nk, ni * nj * (nk + 1))
    for q, (a, nm) in enumerate(zip(arrays, names)):  # every size first, then the pointers
        if a is not None and a.numel() != counts[q % 3]:
            raise ValueError(f"{nm} has {a.numel()} elements, expected {counts[q % 3]}")
    return [_ptr(a, numel=counts[q % 3], name=nm) for q, (a, nm) in enumerate(zip(arrays, names))]


_FACE_NAMES = ("u", "v", "w", "uw", "vw", "ww", "us", "vs", "ws")


def _sizes(*items):
    """(tensor, numel, name) triples: raises ValueError for the first wrong element count."""
    for t, numel, name in items:
        if t is not None and t.numel() != numel:
            raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def default_opts(**kw):
    o = SolveOpts()
    lib().vflip_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Solver:
    """Owns a vflip_solver (MG hierarchy + PCG workspace) for an ni x nj x nk grid."""

    def __init__(self, ni, nj, nk, **opts):
        self.ni, self.nj, self.nk = ni, nj, nk
        self.n = ni * nj * nk
        self.opts = default_opts(**opts)
        h = ctypes.c_void_p()
        check(lib().vflip_solver_create(ni, nj, nk, ctypes.byref(self.opts), ctypes.byref(h)), "solver_create")
        self.h = h
        check(lib().vflip_solver_set_stream(self.h, _stream()), "set_stream")

    def close(self):
        if getattr(self, "h", None):
            lib().vflip_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_opts(self, **kw):
        for k, v in kw.items():
            setattr(self.opts, k, v)
        check(lib().vflip_solver_set_opts(self.h, ctypes.byref(self.opts)), "set_opts")

    def levels(self):
        out = []
        for l in range(lib().vflip_solver_levels(self.h)):
            a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(lib().vflip_solver_level_dims(self.h, l, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
            out.append((a.value, b.value, c.value))
        return out

    def set_matrix(self, d, ax, ay, az, leak=None):
        n = self.n
        check(lib().vflip_solver_set_matrix(self.h, _ptr(d, numel=n), _ptr(ax, numel=n), _ptr(ay, numel=n),
                                            _ptr(az, numel=n), _ptr(leak, numel=n)), "set_matrix")

    def setup(self):
        check(lib().vflip_solver_setup(self.h), "setup")

    def solve(self, rhs, x):
        st = SolveStats()
        rc = check(lib().vflip_solver_solve(self.h, _ptr(rhs, numel=self.n, name="rhs"),
                                            _ptr(x, numel=self.n, name="x"), ctypes.byref(st)), "solve")
        d = st.as_dict()
        d["status"] = rc
        return d

    def solve_f64(self, rhs, x):
        st = SolveStats()
        rc = check(lib().vflip_solver_solve_f64(self.h, _ptr(rhs, torch.float64, self.n, "rhs"),
                                                _ptr(x, torch.float64, self.n, "x"), ctypes.byref(st)), "solve_f64")
        d = st.as_dict()
        d["status"] = rc
        return d

    def set_diag64(self, d):
        check(lib().vflip_solver_set_diag64(self.h, _ptr(d, torch.float64, self.n, "diag")), "set_diag64")

    def vcycle(self, r, z):
        check(lib().vflip_solver_vcycle(self.h, _ptr(r, numel=self.n), _ptr(z, numel=self.n)), "vcycle")

    def matvec(self, x, y):
        check(lib().vflip_solver_matvec(self.h, _ptr(x, numel=self.n), _ptr(y, numel=self.n)), "matvec")

    def profile(self, enable=True):
        check(lib().vflip_solver_profile(self.h, int(enable)), "profile")

    def profile_read(self):
        n, ms, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().vflip_solver_profile_read(self.h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
              "profile_read")
        return {"launches": n.value, "total_ms": ms.value, "bytes_per_launch": b.value}

    def get_level(self, l):
        a, b, c = self.levels()[l]
        outs = [torch.empty((c, b, a), device="cuda", dtype=torch.float32) for _ in range(4)]
        check(lib().vflip_solver_get_level(self.h, l, *[_ptr(o) for o in outs]), "get_level")
        return outs

    def _faces(self, *arrays):
        return _mac_ptrs(self.ni, self.nj, self.nk, arrays, _FACE_NAMES)

    def assemble(self, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs):
        _sizes((rhs, self.n, "rhs"))
        check(lib().vflip_assemble(self.h, variant, h, dt, *self._faces(u, v, w, uw, vw, ww, us, vs, ws),
                                   _ptr(rhs, numel=self.n, name="rhs")), "assemble")

    def assemble_liquid(self, h, dt, u, v, w, uw, vw, ww, us, vs, ws, liquid_phi, rhs):
        """Free-surface assembly: cells with liquid_phi >= 0 are air (Dirichlet p = 0)."""
        _sizes((rhs, self.n, "rhs"), (liquid_phi, self.n, "liquid_phi"))
        check(lib().vflip_assemble_liquid(self.h, h, dt, *self._faces(u, v, w, uw, vw, ww, us, vs, ws),
                                          _ptr(liquid_phi, numel=self.n, name="liquid_phi"),
                                          _ptr(rhs, numel=self.n, name="rhs")), "assemble_liquid")

    def project_learned(self, net, factor, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs, p):
        """project() with the U-Net's pressure guess as the PCG's initial iterate."""
        _sizes((rhs, self.n, "rhs"), (p, self.n, "p"))
        st = SolveStats()
        rc = check(lib().vflip_project_learned(self.h, net.h, factor, variant, h, dt,
                                               *self._faces(u, v, w, uw, vw, ww, us, vs, ws),
                                               _ptr(rhs, numel=self.n, name="rhs"), _ptr(p, numel=self.n, name="p"),
                                               ctypes.byref(st)),
                   "project_learned")
        d = st.as_dict()
        d["status"] = rc
        return d

    def project_history(self, net, factor, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs, p):
        """project() from the solver's pressure history (2 p1 - p2), corrected by the
        network from the divergence the guess leaves when net is given (else None)."""
        _sizes((rhs, self.n, "rhs"), (p, self.n, "p"))
        st = SolveStats()
        rc = check(lib().vflip_project_history(self.h, None if net is None else net.h, factor, variant, h, dt,
                                               *self._faces(u, v, w, uw, vw, ww, us, vs, ws),
                                               _ptr(rhs, numel=self.n, name="rhs"), _ptr(p, numel=self.n, name="p"),
                                               ctypes.byref(st)),
                   "project_history")
        d = st.as_dict()
        d["status"] = rc
        return d

    def history_guess(self, rhs, out=None):
        """The FP64 guess project_history would start from for the current operator and rhs;
        returns (guess, number of history entries used)."""
        out = torch.empty(rhs.shape, device="cuda", dtype=torch.float64) if out is None else out
        used = check(lib().vflip_solver_history_guess(self.h, _ptr(rhs, numel=self.n, name="rhs"),
                                                      _ptr(out, torch.float64, self.n, "guess")), "history_guess")
        return out, used

    def reset_history(self):
        check(lib().vflip_solver_reset_history(self.h), "reset_history")

    def assembled64(self):
        """(diag, rhs) in FP64 as the last assembly wrote them (device tensors)."""
        d = torch.empty((self.nk, self.nj, self.ni), device="cuda", dtype=torch.float64)
        b = torch.empty_like(d)
        check(lib().vflip_solver_get_assembled64(self.h, _ptr(d, torch.float64), _ptr(b, torch.float64)),
              "get_assembled64")
        return d, b

    def solve_refined(self, rhs, x, tol=1e-9, max_rounds=8, assembled=True, x64=None):
        """FP64 defect correction around this solver's own solves (vflip_solver_solve_refined)."""
        _sizes((rhs, self.n, "rhs"), (x, self.n, "x"), (x64, self.n, "x64"))
        st = SolveStats()
        rc = check(lib().vflip_solver_solve_refined(self.h, _ptr(rhs, numel=self.n, name="rhs"),
                                                    _ptr(x, numel=self.n, name="x"),
                                                    _ptr(x64, torch.float64, self.n, "x64"), int(assembled),
                                                    float(tol), int(max_rounds), ctypes.byref(st)), "solve_refined")
        d = st.as_dict()
        d["status"] = rc
        return d

    def project(self, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs, p):
        _sizes((rhs, self.n, "rhs"), (p, self.n, "p"))
        st = SolveStats()
        rc = check(lib().vflip_project(self.h, variant, h, dt, *self._faces(u, v, w, uw, vw, ww, us, vs, ws),
                                       _ptr(rhs, numel=self.n, name="rhs"), _ptr(p, numel=self.n, name="p"),
                                       ctypes.byref(st)),
                   "project")
        d = st.as_dict()
        d["status"] = rc
        return d


def face_weights(phi):
    nk, nj, ni = (s - 1 for s in phi.shape)
    uw = torch.empty((nk, nj, ni + 1), device=phi.device, dtype=torch.float32)
    vw = torch.empty((nk, nj + 1, ni), device=phi.device, dtype=torch.float32)
    ww = torch.empty((nk + 1, nj, ni), device=phi.device, dtype=torch.float32)
    check(lib().vflip_face_weights(ni, nj, nk, _ptr(phi), _ptr(uw), _ptr(vw), _ptr(ww), _stream()), "face_weights")
    return uw, vw, ww


def inflow_bc(bdesc, u_in, u, v, w):
    nk, nj, ni = bdesc.shape
    faces = _mac_ptrs(ni, nj, nk, (u, v, w), ("u", "v", "w"))
    check(lib().vflip_inflow_bc(ni, nj, nk, _ptr(bdesc, torch.int8, name="bdesc"), u_in, *faces, _stream()),
          "inflow_bc")


def grad_apply(variant, h, dt, p, uw, vw, ww, us, vs, ws, u, v, w):
    nk, nj, ni = p.shape
    faces = _mac_ptrs(ni, nj, nk, (uw, vw, ww, us, vs, ws, u, v, w), _FACE_NAMES[3:] + _FACE_NAMES[:3])
    check(lib().vflip_grad_apply(ni, nj, nk, variant, h, dt, _ptr(p, name="p"), *faces, _stream()), "grad_apply")


def grad_apply_f64(variant, h, dt, p, uw, vw, ww, us, vs, ws, u, v, w):
    """grad_apply from an FP64 pressure ((float)(p[c] - p[c-1]) as the reference)."""
    nk, nj, ni = p.shape
    faces = _mac_ptrs(ni, nj, nk, (uw, vw, ww, us, vs, ws, u, v, w), _FACE_NAMES[3:] + _FACE_NAMES[:3])
    check(lib().vflip_grad_apply_f64(ni, nj, nk, variant, h, dt, _ptr(p, torch.float64, name="p"), *faces,
                                     _stream()), "grad_apply_f64")


def set_boundary_phi(ni, nj, nk, h, fn, out=None):
    """set_boundary_phi: nodal (nk+1, nj+1, ni+1) device array of fn(x, y, z) (a Python
    callable taking three floats) at the reference's node positions."""
    out = torch.empty((nk + 1, nj + 1, ni + 1), device="cuda") if out is None else out
    cb = PHI_FN(lambda pos, user: float(fn(pos[0], pos[1], pos[2])))
    check(lib().vflip_set_boundary_phi(ni, nj, nk, h, ctypes.cast(cb, ctypes.c_void_p), None,
                                       _ptr(out, numel=(ni + 1) * (nj + 1) * (nk + 1), name="nodal_phi"), _stream()),
          "set_boundary_phi")
    return out


def set_boundary_vel(ni, nj, nk, h, fn):
    """set_boundary_vel: MAC solid velocity (us, vs, ws) from fn(x, y, z) -> (vx, vy, vz)."""
    us = torch.empty((nk, nj, ni + 1), device="cuda")
    vs = torch.empty((nk, nj + 1, ni), device="cuda")
    ws = torch.empty((nk + 1, nj, ni), device="cuda")

    def call(pos, out, user):
        r = fn(pos[0], pos[1], pos[2])
        out[0], out[1], out[2] = float(r[0]), float(r[1]), float(r[2])
    cb = VEL_FN(call)
    check(lib().vflip_set_boundary_vel(ni, nj, nk, h, ctypes.cast(cb, ctypes.c_void_p), None,
                                       *_mac_ptrs(ni, nj, nk, (us, vs, ws), ("us", "vs", "ws")), _stream()),
          "set_boundary_vel")
    return us, vs, ws


def amgpcg_solve_host(diag, ax, ay, az, rhs, tol=1e-6, maxit=50, precision=MIXED):
    """AMGPCGSolve drop-in on host float64 numpy arrays shaped (nk, nj, ni)."""
    import numpy as np
    nk, nj, ni = diag.shape
    arrs = [np.ascontiguousarray(a, np.float64) for a in (diag, ax, ay, az, rhs)]
    res = np.zeros_like(arrs[0])
    r = ctypes.c_double()
    it = ctypes.c_int()
    o = default_opts(precision=precision)
    rc = check(lib().vflip_amgpcg_solve_host(ni, nj, nk, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs],
                                             res.ctypes.data_as(ctypes.c_void_p), tol, maxit, ctypes.byref(r),
                                             ctypes.byref(it), ctypes.byref(o)), "amgpcg_solve_host")
    return rc, res, r.value, it.value


def advect_rk2(h, dt, u, v, w, field, offs, out, mode=0, field2=None, out2=None):
    """RK2 semi-Lagrangian advection of field (and field2) on a grid shaped like field."""
    nk, nj, ni1 = u.shape
    fz, fy, fx = field.shape
    nf = field.numel()
    for t, nm in ((out, "out"), (field2, "field2"), (out2, "out2")):
        if t is not None and t.numel() != nf:
            raise ValueError(f"{nm} has {t.numel()} elements, expected {nf}")
    vel = _mac_ptrs(ni1 - 1, nj, nk, (u, v, w), ("u", "v", "w"))
    check(lib().vflip_advect_rk2(ni1 - 1, nj, nk, h, dt, mode, *vel, _ptr(field, name="field"), fx, fy, fz,
                                 offs[0], offs[1], offs[2], _ptr(out, numel=nf, name="out"),
                                 _ptr(field2, numel=nf, name="field2"), _ptr(out2, numel=nf, name="out2"), _stream()),
          "advect_rk2")


def advect_plume(h, dt, u, v, w, rho, temp, u2, v2, w2, rho2, temp2):
    """The smoke step's advection of u, v, w, rho and T in one launch (mode 2 arithmetic)."""
    nk, nj, ni = rho.shape
    faces = _mac_ptrs(ni, nj, nk, (u, v, w, u2, v2, w2), ("u", "v", "w", "u2", "v2", "w2"))
    cells = [_ptr(t, numel=ni * nj * nk, name=nm) for t, nm in ((rho, "rho"), (temp, "temp"), (rho2, "rho2"),
                                                                (temp2, "temp2"))]
    check(lib().vflip_advect_plume(ni, nj, nk, h, dt, *faces[:3], cells[0], cells[1], *faces[3:], cells[2], cells[3],
                                   _stream()), "advect_plume")


def plume_force(h, dt, center, r, seed, jitter, alpha, beta, rho, temp, v):
    nk, nj, ni = rho.shape
    check(lib().vflip_plume_force(ni, nj, nk, h, dt, center[0], center[1], center[2], r, seed, jitter, alpha, beta,
                                  _ptr(rho), _ptr(temp), _ptr(v), _stream()), "plume_force")


class Batch:
    """Batched pressure solves of several FP32 subdomain solvers (vflip_batch): every
    V-cycle / PCG step is one launch over all subdomains, the stop tests run on the
    GPU and a fixed number of iterations replays from a captured hipGraph."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = ctypes.c_void_p()
        check(lib().vflip_batch_create(len(self.solvers), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(h)),
              "batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vflip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph_iterations(self, k):
        check(lib().vflip_batch_set_graph_iterations(self.h, k), "batch_set_graph_iterations")

    def autotune(self, levels=8):
        """Times the fused sweep's tile shapes per level depth; returns [(rows, z chunk)] per depth."""
        ty, zc = (ctypes.c_int * levels)(), (ctypes.c_int * levels)()
        check(lib().vflip_batch_autotune(self.h, ctypes.cast(ty, ctypes.c_void_p), ctypes.cast(zc, ctypes.c_void_p),
                                         levels), "batch_autotune")
        return [(ty[i], zc[i]) for i in range(levels)]

    def solve(self, rhs, x, graph=True):
        """rhs, x: lists of device FP32 tensors (one per solver). Returns (status, [stats], ms)."""
        n = len(self.solvers)
        assert len(rhs) == n and len(x) == n
        r = (ctypes.c_void_p * n)(*[_ptr(t, numel=s.n, name="rhs").value for t, s in zip(rhs, self.solvers)])
        xs = (ctypes.c_void_p * n)(*[_ptr(t, numel=s.n, name="x").value for t, s in zip(x, self.solvers)])
        st = (SolveStats * n)()
        ms = ctypes.c_double()
        rc = check(lib().vflip_batch_solve(self.h, ctypes.cast(r, ctypes.c_void_p), ctypes.cast(xs, ctypes.c_void_p),
                                           ctypes.cast(st, ctypes.c_void_p), int(graph), ctypes.byref(ms)),
                   "batch_solve")
        out = []
        for i in range(n):
            d = st[i].as_dict()
            out.append(d)
        return rc, out, ms.value

    def solve_refined(self, rhs, x, tol=1e-9, max_rounds=8, assembled=True, x64=None, graph=True):
        """FP64 defect correction around the batched FP32 solve (vflip_batch_solve_refined):
        every subdomain to an FP64 true residual of tol. Returns (status, [stats], ms)."""
        n = len(self.solvers)
        assert len(rhs) == n and len(x) == n and (x64 is None or len(x64) == n)
        r = (ctypes.c_void_p * n)(*[_ptr(t, numel=s.n, name="rhs").value for t, s in zip(rhs, self.solvers)])
        xs = (ctypes.c_void_p * n)(*[_ptr(t, numel=s.n, name="x").value for t, s in zip(x, self.solvers)])
        x64s = None if x64 is None else (ctypes.c_void_p * n)(
            *[_ptr(t, torch.float64, s.n, "x64").value for t, s in zip(x64, self.solvers)])
        st = (SolveStats * n)()
        ms = ctypes.c_double()
        rc = check(lib().vflip_batch_solve_refined(self.h, ctypes.cast(r, ctypes.c_void_p),
                                                   ctypes.cast(xs, ctypes.c_void_p),
                                                   None if x64s is None else ctypes.cast(x64s, ctypes.c_void_p),
                                                   int(assembled), float(tol), int(max_rounds),
                                                   ctypes.cast(st, ctypes.c_void_p), int(graph), ctypes.byref(ms)),
                   "batch_solve_refined")
        return rc, [st[i].as_dict() for i in range(n)], ms.value


# ---------------------------------------------------------------------- FLIP
class Particles:
    """Binning/transfer workspace (vflip_particles) for an ni x nj x nk grid of spacing h.

    Particle arrays are device float32 tensors shaped (np, 3); grids are (z, y, x)
    tensors in the MAC shapes u:(nk, nj, ni+1), v:(nk, nj+1, ni), w:(nk+1, nj, ni).
    """

    def __init__(self, ni, nj, nk, h):
        self.ni, self.nj, self.nk, self.h = ni, nj, nk, h
        self.ncell = ni * nj * nk
        hnd = ctypes.c_void_p()
        check(lib().vflip_particles_create(ni, nj, nk, h, ctypes.byref(hnd)), "particles_create")
        self.hnd = hnd
        check(lib().vflip_particles_set_stream(self.hnd, _stream()), "particles_set_stream")

    def close(self):
        if getattr(self, "hnd", None):
            lib().vflip_particles_destroy(self.hnd)
            self.hnd = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hash(self, pos, clamp=False, bmin=(0.0, 0.0, 0.0), count=True):
        """hash_FLIP: returns (cell_key[np], cell_start[ncell+1], order[np], nbinned or None)."""
        npart = pos.shape[0]
        key = torch.empty(max(npart, 1), dtype=torch.int32, device="cuda")
        start = torch.empty(self.ncell + 1, dtype=torch.int32, device="cuda")
        order = torch.empty(max(npart, 1), dtype=torch.int32, device="cuda")
        nb = ctypes.c_int(0)
        check(lib().vflip_hash(self.hnd, npart, _ptr(pos, numel=3 * npart, name="pos") if npart else None,
                               int(clamp), *[float(b) for b in bmin], _ptr(key, torch.int32),
                               _ptr(start, torch.int32), _ptr(order, torch.int32),
                               ctypes.byref(nb) if count else None), "hash")
        return key[:npart], start, order, (nb.value if count else None)

    def p2g(self, pos, vel, start, order, variant=COARSE, method=0, bmin=(0.0, 0.0, 0.0), out=None):
        """particle_to_grid for u, v, w: returns [(u, u_coef), (v, v_coef), (w, w_coef)]."""
        ni, nj, nk = self.ni, self.nj, self.nk
        if out is None:
            shapes = [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]
            out = [(torch.empty(s, device="cuda"), torch.empty(s, device="cuda")) for s in shapes]
        npart = pos.shape[0]
        ptrs = []
        for f, c in out:
            ptrs += [_ptr(f), _ptr(c)]
        check(lib().vflip_p2g(self.hnd, variant, method, npart, _ptr(pos) if npart else None,
                              _ptr(vel) if npart else None, _ptr(start, torch.int32),
                              _ptr(order, torch.int32), *[float(b) for b in bmin], *ptrs), "p2g")
        return out


    def reorder(self, pos, vel, start, order, u, v, w, solid_phi, seed, variant=COARSE, bmin=(0.0, 0.0, 0.0),
                out=None):
        """reorder_FLIP over hash()'s lists: returns the new (pos, vel), (n, 3) each.

        out = (pos_buf, vel_buf) of shape (cap, 3) is reused when it fits; the
        result is then a view of it. Otherwise buffers of the exact size are made.
        """
        npart = pos.shape[0]
        if out is None:
            cap = npart + npart // 8 + 1024
            out = (torch.empty((cap, 3), device="cuda"), torch.empty((cap, 3), device="cuda"))
        for attempt in range(2):
            cap = min(out[0].shape[0], out[1].shape[0])
            n = ctypes.c_int(0)
            rc = lib().vflip_reorder(self.hnd, variant, *[float(b) for b in bmin], int(seed) & 0xFFFFFFFF, npart,
                                     _ptr(pos) if npart else None, _ptr(vel) if npart else None,
                                     _ptr(start, torch.int32, numel=self.ncell + 1),
                                     _ptr(order, torch.int32) if npart else None, _ptr(u), _ptr(v), _ptr(w),
                                     _ptr(solid_phi), cap, _ptr(out[0]) if cap else None,
                                     _ptr(out[1]) if cap else None, ctypes.byref(n))
            if rc == ERR_CAPACITY and attempt == 0:
                out = (torch.empty((n.value, 3), device="cuda"), torch.empty((n.value, 3), device="cuda"))
                continue
            check(rc, "reorder")
            return out[0][:n.value], out[1][:n.value]


def nested_views(coarse, sub):
    """ctypes views of the coarse grids and one subdomain.

    coarse / sub: dicts of device float32 tensors u, v, w (MAC shapes), phi (nodal)
    and optional us, vs, ws, plus h; sub also bmin, bmax (3 floats each)."""
    def ptr(d, k):
        t = d.get(k)
        return None if t is None else _ptr(t, name=k)

    nk, nj, ni1 = coarse["u"].shape
    C = CoarseView(ni1 - 1, nj, nk, float(coarse["h"]), *[ptr(coarse, k) for k in ("u", "v", "w", "phi", "us", "vs", "ws")])
    snk, snj, sni1 = sub["u"].shape
    S = SubdomainView(sni1 - 1, snj, snk, float(sub["h"]), (ctypes.c_float * 3)(*sub["bmin"]),
                      (ctypes.c_float * 3)(*sub["bmax"]), *[ptr(sub, k) for k in ("u", "v", "w", "phi", "us", "vs", "ws")])
    return C, S


def nested_velocity(coarse, sub, pos):
    """get_velocity with the subdomain blend at pos (n, 3); returns (n, 3)."""
    C, S = nested_views(coarse, sub)
    out = torch.empty_like(pos)
    n = pos.shape[0]
    check(lib().vflip_nested_velocity(ctypes.byref(C), ctypes.byref(S), n, _ptr(pos) if n else None,
                                      _ptr(out) if n else None, _stream()), "nested_velocity")
    return out


def advect_particles_nested(coarse, sub, dt, pos):
    """advect_FLIP of the coarse particles with a subdomain, in place on pos."""
    C, S = nested_views(coarse, sub)
    n = pos.shape[0]
    check(lib().vflip_advect_particles_nested(ctypes.byref(C), ctypes.byref(S), float(dt), n,
                                              _ptr(pos) if n else None, _stream()), "advect_particles_nested")
    return pos


def pse_diffuse(coarse, sub, dt, nu, np_real, pos, vel, start, order, out=None):
    """diffuse (PSE viscosity): new velocities of the first np_real subdomain particles."""
    C, S = nested_views(coarse, sub)
    out = torch.empty((np_real, 3), device="cuda") if out is None else out
    check(lib().vflip_pse_diffuse(ctypes.byref(C), ctypes.byref(S), float(dt), float(nu), np_real,
                                  _ptr(pos), _ptr(vel), _ptr(start, torch.int32), _ptr(order, torch.int32),
                                  _ptr(out) if np_real else None, _stream()), "pse_diffuse")
    return out


def _selection(fn, what, cap0):
    """Runs a synchronous selection entry point fn(cap, out_pos, out_vel, nout), growing the output once."""
    cap = cap0
    for attempt in range(2):
        op = torch.empty((max(cap, 1), 3), device="cuda")
        ov = torch.empty((max(cap, 1), 3), device="cuda")
        n = ctypes.c_int(0)
        rc = fn(cap, _ptr(op), _ptr(ov), ctypes.byref(n))
        if rc == ERR_CAPACITY and attempt == 0:
            cap = n.value
            continue
        check(rc, what)
        return op[:n.value], ov[:n.value]


def halo_select(P, sub, dx, pos, vel):
    """The particles with is_halo(pos, dx), in order: (pos, vel)."""
    _, S = nested_views(sub, sub)
    n = pos.shape[0]
    return _selection(lambda cap, a, b, c: lib().vflip_halo_select(P.hnd, ctypes.byref(S), float(dx), n,
                                                                   _ptr(pos) if n else None,
                                                                   _ptr(vel) if n else None, cap, a, b, c),
                      "halo_select", n)


def seed_ghost(P, coarse, sub, seed):
    """seed_ghost over the subdomain (P: its workspace): (pos, vel) of the ghosts."""
    C, S = nested_views(coarse, sub)
    return _selection(lambda cap, a, b, c: lib().vflip_seed_ghost(P.hnd, ctypes.byref(C), ctypes.byref(S),
                                                                  int(seed) & 0xFFFFFFFF, cap, a, b, c),
                      "seed_ghost", 2 * S.ni * S.nj * S.nk)


def advect_tracers(coarse, sub, dt, decay, tracers):
    """advect_tracers in place on a (n, 4) float32 tensor; sub None = no subdomain."""
    if sub is None:
        C, _ = nested_views(coarse, dict(coarse, bmin=(0, 0, 0), bmax=(0, 0, 0)))
        sref = None
    else:
        C, S = nested_views(coarse, sub)
        sref = ctypes.byref(S)
    n = tracers.shape[0]
    check(lib().vflip_advect_tracers(ctypes.byref(C), sref, float(dt), float(decay), n,
                                     _ptr(tracers) if n else None, _stream()), "advect_tracers")
    return tracers


def emit_tracers_sphere(P, seed, n_emit, centre, r, hx):
    out = torch.empty((max(n_emit, 1), 4), device="cuda")
    check(lib().vflip_emit_tracers_sphere(P.hnd, int(seed) & 0xFFFFFFFF, n_emit, *[float(c) for c in centre],
                                          float(r), float(hx), _ptr(out)), "emit_tracers_sphere")
    return out[:n_emit]


def emit_tracers_sub(P, coarse, sub, seed):
    """Subdomain tracers near the solid (P: the subdomain's workspace): (n, 4)."""
    C, S = nested_views(coarse, sub)
    cap = 4 * S.ni * S.nj * S.nk
    out = torch.empty((max(cap, 1), 4), device="cuda")
    n = ctypes.c_int(0)
    check(lib().vflip_emit_tracers_sub(P.hnd, ctypes.byref(C), ctypes.byref(S), int(seed) & 0xFFFFFFFF, cap,
                                       _ptr(out), ctypes.byref(n)), "emit_tracers_sub")
    return out[:n.value]


def tracer_frame(tracers, lx):
    out = torch.empty_like(tracers)
    n = tracers.shape[0]
    check(lib().vflip_tracer_frame(n, _ptr(tracers) if n else None, float(lx), _ptr(out) if n else None, _stream()),
          "tracer_frame")
    return out


def write_tracers(directory, frame, tracers, lx):
    """write_tracers: directory/Particle_data%04d.bin, the frame records as raw float32."""
    rec = tracer_frame(tracers, lx).cpu().numpy().astype("<f4")
    path = os.path.join(directory, "Particle_data%04d.bin" % frame)
    with open(path, "wb") as f:
        f.write(rec.tobytes())
    return path


def liquid_phi(h, radius, pos, cell_start, order, shape, out=None):
    """Particle level set at the cell centres of a grid shaped (nk, nj, ni) from vflip_hash's lists."""
    nk, nj, ni = shape
    out = torch.empty(shape, device="cuda") if out is None else out
    npart = pos.shape[0]
    check(lib().vflip_liquid_phi(ni, nj, nk, h, radius, _ptr(pos) if npart else None, _ptr(cell_start, torch.int32),
                                 _ptr(order, torch.int32), _ptr(out, numel=ni * nj * nk), _stream()), "liquid_phi")
    return out


def get_delta(new, old, out=None):
    out = torch.empty_like(new) if out is None else out
    check(lib().vflip_get_delta(new.numel(), _ptr(new), _ptr(old, numel=new.numel()), _ptr(out, numel=new.numel()),
                                _stream()), "get_delta")
    return out


def g2p(h, pos, vel, new, old=None, flip=1.0, bmin=(0.0, 0.0, 0.0)):
    """particle_from_grid in place on vel: new = (u, v, w) grids, old = (u, v, w) or None (new are deltas)."""
    nk, nj, ni1 = new[0].shape
    npart = pos.shape[0]
    olds = [None, None, None] if old is None else [_ptr(o) for o in old]
    check(lib().vflip_g2p(ni1 - 1, nj, nk, h, npart, _ptr(pos) if npart else None, _ptr(vel) if npart else None,
                          *[_ptr(g) for g in new], *olds, flip, *[float(b) for b in bmin], _stream()), "g2p")


def advect_particles(h, dt, pos, u, v, w, solid_phi, variant=COARSE, bmin=(0.0, 0.0, 0.0)):
    """advect_FLIP in place on pos (traceRK3 + solid push-out)."""
    nk, nj, ni1 = u.shape
    npart = pos.shape[0]
    check(lib().vflip_advect_particles(ni1 - 1, nj, nk, h, dt, variant, *[float(b) for b in bmin], npart,
                                       _ptr(pos) if npart else None, _ptr(u), _ptr(v), _ptr(w), _ptr(solid_phi),
                                       _stream()), "advect_particles")


# ------------------------------------------------------ coarse/fine coupling
def fine_to_coarse(fine, coarse, ratio, offs=(0, 0, 0)):
    fz, fy, fx = fine.shape
    cz, cy, cx = coarse.shape
    check(lib().vflip_fine_to_coarse(_ptr(fine), fx, fy, fz, _ptr(coarse), cx, cy, cz, ratio, *offs, _stream()),
          "fine_to_coarse")


def coarse_to_fine(fine, f_offs, hf, coarse, c_offs, hc, fine_minus=None):
    """fine = coarse sampled at fine positions; with fine_minus: fine = fine_minus - sample."""
    fz, fy, fx = fine.shape
    cz, cy, cx = coarse.shape
    check(lib().vflip_coarse_to_fine(_ptr(fine), fx, fy, fz, *f_offs, hf, _ptr(coarse), cx, cy, cz, *c_offs, hc,
                                     _ptr(fine_minus, numel=fine.numel()), _stream()), "coarse_to_fine")


def set_face_boundary(fine, coarse, ratio, offs):
    fz, fy, fx = fine.shape
    cz, cy, cx = coarse.shape
    check(lib().vflip_set_face_boundary(_ptr(fine), fx, fy, fz, _ptr(coarse), cx, cy, cz, ratio, *offs, _stream()),
          "set_face_boundary")


def to_bricks(src):
    nz, ny, nx = src.shape
    dst = torch.empty(((nx + 7) // 8) * ((ny + 7) // 8) * ((nz + 7) // 8) * 512, device=src.device)
    check(lib().vflip_to_bricks(_ptr(src), nx, ny, nz, _ptr(dst), _stream()), "to_bricks")
    return dst


def from_bricks(src, shape):
    nz, ny, nx = shape
    dst = torch.empty(shape, device=src.device)
    check(lib().vflip_from_bricks(_ptr(src), nx, ny, nz, _ptr(dst), _stream()), "from_bricks")
    return dst


# ------------------------------------------------------ geometric multigrid (MultiGridSolver3D<T>)
_MG3D_TYPES = {torch.float32: 0, torch.float64: 1}


def _mg3d_type(t, name="x"):
    if t.dtype not in _MG3D_TYPES:
        raise ValueError(f"{name} must be float32 or float64, got {t.dtype}")
    return _MG3D_TYPES[t.dtype]


class MG3D:
    """MultiGridSolver3D<T> (S/Multigrid3D.h:152-184) on an (nz, ny, nx) grid, T = dtype
    (float32: MGSolverf, float64: MGSolverd); dirichlet is the reference's _b_Dirichlet."""

    def __init__(self, nx, ny, nz, dtype=torch.float32, dirichlet=False):
        if dtype not in _MG3D_TYPES:
            raise ValueError(f"dtype must be float32 or float64, got {dtype}")
        self.shape, self.dtype, self.dirichlet = (nz, ny, nx), dtype, bool(dirichlet)
        h = ctypes.c_void_p()
        check(lib().vflip_mg3d_create(nx, ny, nz, _MG3D_TYPES[dtype], int(dirichlet), ctypes.byref(h)), "mg3d_create")
        self.h = h

    def close(self):
        h = getattr(self, "h", None)
        if h:
            lib().vflip_mg3d_destroy(h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def levels(self):
        """[(nx, ny, nz)] of every level, finest first (m_ComputeLevels / systemk)."""
        n = ctypes.c_int()
        check(lib().vflip_mg3d_levels(self.h, ctypes.byref(n), None), "mg3d_levels")
        dims = (ctypes.c_int * (3 * n.value))()
        check(lib().vflip_mg3d_levels(self.h, ctypes.byref(n), dims), "mg3d_levels")
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(n.value)]

    def _args(self, x, b):
        n = self.shape[0] * self.shape[1] * self.shape[2]
        _sizes((x, n, "x"), (b, n, "b"))
        return _ptr(x, self.dtype, n, "x"), _ptr(b, self.dtype, n, "b")

    def vcycle(self, x, b):
        """m_Vcycle(x, b, tol, residual, 0): one V-cycle on x in place."""
        check(lib().vflip_mg3d_vcycle(self.h, *self._args(x, b), _stream()), "mg3d_vcycle")

    def full_multigrid(self, x, b):
        """m_FullMultiGrid(x, b, tol, residual) on x in place."""
        check(lib().vflip_mg3d_full_multigrid(self.h, *self._args(x, b), _stream()), "mg3d_full_multigrid")


def mg3d_rbgs(b, x, iters, omega=1.0, dirichlet=False, x_new=None):
    """m_Red_Black_Gauss on x (nz, ny, nx) in place; x_new is scratch of x's shape."""
    ty = _mg3d_type(x)
    nz, ny, nx = x.shape
    if x_new is None:
        x_new = torch.empty_like(x)
    _sizes((b, x.numel(), "b"), (x_new, x.numel(), "x_new"))
    check(lib().vflip_mg3d_rbgs(nx, ny, nz, ty, int(dirichlet), _ptr(b, x.dtype, name="b"), _ptr(x, x.dtype),
                                _ptr(x_new, x.dtype, name="x_new"), int(iters), float(omega), _stream()), "mg3d_rbgs")


def mg3d_residual(b, x, dirichlet=False, out=None):
    """m_ComputeResidual: b - A x."""
    ty = _mg3d_type(x)
    nz, ny, nx = x.shape
    out = torch.empty_like(x) if out is None else out
    _sizes((b, x.numel(), "b"), (out, x.numel(), "out"))
    check(lib().vflip_mg3d_residual(nx, ny, nz, ty, int(dirichlet), _ptr(out, x.dtype, name="out"),
                                    _ptr(b, x.dtype, name="b"), _ptr(x, x.dtype), _stream()), "mg3d_residual")
    return out


def mg3d_coarse_shape(shape):
    return tuple(max(n // 2, 1) for n in shape)


def mg3d_restrict(src):
    """m_Restrict: the coarser grid's 0.5 * sum of 8 children."""
    ty = _mg3d_type(src, "src")
    nz, ny, nx = src.shape
    dst = torch.empty(mg3d_coarse_shape(src.shape), dtype=src.dtype, device=src.device)
    check(lib().vflip_mg3d_restrict(nx, ny, nz, ty, _ptr(src, src.dtype, name="src"), _ptr(dst, src.dtype, name="dst"),
                                    _stream()), "mg3d_restrict")
    return dst


def mg3d_prolongate(src, dst):
    """m_Prolongate: dst += parent values from the coarser grid src (in place on dst)."""
    ty = _mg3d_type(dst, "dst")
    nz, ny, nx = dst.shape
    cshape = mg3d_coarse_shape(dst.shape)
    if tuple(src.shape) != cshape:
        raise ValueError(f"src has shape {tuple(src.shape)}, expected {cshape}")
    check(lib().vflip_mg3d_prolongate(nx, ny, nz, ty, _ptr(src, dst.dtype, name="src"),
                                      _ptr(dst, dst.dtype, name="dst"), _stream()), "mg3d_prolongate")
