"""Smoke-plume time step on the GPU, driven through the C ABI.

One step, on the reference's MAC grid (cell (i,j,k) centre at (i,j,k)h):
  1. RK2 semi-Lagrangian advection of u, v, w, rho and T by the current
     velocity (SubDomainSolver::advect_field semantics, S/sub_domain.cpp:191-235),
     one launch for the five fields;
  2. sphere source (rho = T = 1 + jitter) and face buoyancy
     v += dt(-alpha rho + beta T);
  3. pressure projection with static cut-cell face weights
     (SmokeSolver3D::pcg_projection, S/Smoke_solver3D.cpp:945-1200): assembly,
     MGPCG to the relative tolerance, gradient update.
The solver hierarchy is rebuilt only when the assembled operator changes (the
obstacles are static, so after the first step it is reused).
"""
import torch

from . import api


class PlumeSim:
    def __init__(self, n, phi, dt=0.01, precision=api.FP32, tol=1e-6, maxit=50, source_center_cfg=(0.5, 0.1, 0.5),
                 source_r=0.08, jitter=0.05, alpha=0.1, beta=0.3, seed=0, mode=2):
        self.n = n
        self.h = 1.0 / n
        self.dt = dt
        self.mode = mode
        self.fused = True  # mode 2: the five advections in one launch (vflip_advect_plume)
        dev = torch.device("cuda")
        f32 = dict(device=dev, dtype=torch.float32)
        self.u = torch.zeros((n, n, n + 1), **f32)
        self.v = torch.zeros((n, n + 1, n), **f32)
        self.w = torch.zeros((n + 1, n, n), **f32)
        self.u2, self.v2, self.w2 = torch.zeros_like(self.u), torch.zeros_like(self.v), torch.zeros_like(self.w)
        self.rho = torch.zeros((n, n, n), **f32)
        self.temp = torch.zeros_like(self.rho)
        self.rho2, self.temp2 = torch.zeros_like(self.rho), torch.zeros_like(self.rho)
        self.us, self.vs, self.ws = torch.zeros_like(self.u), torch.zeros_like(self.v), torch.zeros_like(self.w)
        self.p = torch.zeros_like(self.rho)
        self.rhs = torch.zeros_like(self.rho)
        self.phi = phi if isinstance(phi, torch.Tensor) else torch.from_numpy(phi).to(dev)
        self.uw, self.vw, self.ww = api.face_weights(self.phi)
        # config coordinates (cell centres at (i+.5)h) -> reference coordinates
        c = [x - 0.5 * self.h for x in source_center_cfg]
        self.source = (c, source_r)
        self.jitter, self.alpha, self.beta, self.seed = jitter, alpha, beta, seed
        self.solver = api.Solver(n, n, n, precision=precision, tolerance=tol, max_iterations=maxit)
        self.last_stats = None
        self.net, self.net_factor = None, 1
        self.history = False

    def use_network(self, net, factor, history=False):
        """Learned warm start: the U-Net (on the grid coarsened by factor) seeds every pressure
        solve; with history, it corrects the pressure-history extrapolation instead."""
        self.net, self.net_factor = net, factor
        self.history = history

    def use_history(self, on=True):
        """Warm start from the pressure history alone (2 p1 - p2), no network."""
        self.history = on
        self.net = None

    def advect(self):
        h, dt = float(self.h), float(self.dt)
        if self.mode == 2 and self.fused:
            api.advect_plume(h, dt, self.u, self.v, self.w, self.rho, self.temp, self.u2, self.v2, self.w2, self.rho2,
                             self.temp2)
        else:
            self._advect_separate(h, dt)
        self.u, self.u2 = self.u2, self.u
        self.v, self.v2 = self.v2, self.v
        self.w, self.w2 = self.w2, self.w
        self.rho, self.rho2 = self.rho2, self.rho
        self.temp, self.temp2 = self.temp2, self.temp

    def _advect_separate(self, h, dt):
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.u, (0.5, 0.0, 0.0), self.u2, mode=self.mode)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.v, (0.0, 0.5, 0.0), self.v2, mode=self.mode)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.w, (0.0, 0.0, 0.5), self.w2, mode=self.mode)
        api.advect_rk2(h, dt, self.u, self.v, self.w, self.rho, (0.0, 0.0, 0.0), self.rho2, mode=self.mode,
                       field2=self.temp, out2=self.temp2)

    def force(self):
        c, r = self.source
        api.plume_force(float(self.h), float(self.dt), c, r, self.seed, self.jitter, self.alpha, self.beta, self.rho,
                        self.temp, self.v)

    def project(self):
        args = (api.COARSE, float(self.h), 1.0, self.u, self.v, self.w, self.uw, self.vw, self.ww, self.us, self.vs,
                self.ws, self.rhs, self.p)
        if self.history:
            self.last_stats = self.solver.project_history(self.net, self.net_factor, *args)
        elif self.net is not None:
            self.last_stats = self.solver.project_learned(self.net, self.net_factor, *args)
        else:
            self.last_stats = self.solver.project(*args)
        if self.last_stats["status"] < 0:
            raise RuntimeError(f"projection failed: {self.last_stats}")
        return self.last_stats

    def step(self):
        self.advect()
        self.force()
        return self.project()

    def state(self):
        return {k: getattr(self, k) for k in ("u", "v", "w", "rho", "temp", "p")}

    def load_state(self, st):
        for k, t in st.items():
            getattr(self, k).copy_(t)
