"""ctypes loader for libvflip.so, declaring every entry point of include/vflip.h.

The library is the product path: there is no CPU fallback. Loading fails with
an error when the HIP build is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvflip.so")

c_int, c_float, c_double, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
P = ctypes.c_void_p  # device or host pointer
PI = ctypes.POINTER(ctypes.c_int)
PD = ctypes.POINTER(ctypes.c_double)


class SolveOpts(ctypes.Structure):
    _fields_ = [
        ("tolerance", c_double),
        ("max_iterations", c_int),
        ("precision", c_int),
        ("pre_sweeps", c_int),
        ("post_sweeps", c_int),
        ("coarse_sweeps", c_int),
        ("refine_every", c_int),
        ("warm_start", c_int),
        ("half_from_level", c_int),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int),
        ("converged", c_int),
        ("residual", c_double),
        ("rel_residual", c_double),
        ("true_residual", c_double),
        ("ms_setup", c_double),
        ("ms_solve", c_double),
        ("warm_rejected", c_int),
        ("guess_residual", c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CoarseView(ctypes.Structure):
    """vflip_coarse_view."""
    _fields_ = [("ni", c_int), ("nj", c_int), ("nk", c_int), ("h", c_float)] + \
        [(f, c_void_p) for f in ("u", "v", "w", "solid_phi", "us", "vs", "ws")]


class SubdomainView(ctypes.Structure):
    """vflip_subdomain_view."""
    _fields_ = [("ni", c_int), ("nj", c_int), ("nk", c_int), ("h", c_float), ("bmin", c_float * 3),
                ("bmax", c_float * 3)] + [(f, c_void_p) for f in ("u", "v", "w", "solid_phi", "us", "vs", "ws")]


# host callbacks of vflip_set_boundary_phi / vflip_set_boundary_vel
PHI_FN = ctypes.CFUNCTYPE(c_float, ctypes.POINTER(c_float), c_void_p)
VEL_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(c_float), ctypes.POINTER(c_float), c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "vflip_last_error": (ctypes.c_char_p, []),
    "vflip_version": (ctypes.c_char_p, []),
    "vflip_device_count": (c_int, []),
    "vflip_default_opts": (None, [ctypes.POINTER(SolveOpts)]),
    "vflip_solver_create": (c_int, [c_int, c_int, c_int, ctypes.POINTER(SolveOpts), ctypes.POINTER(c_void_p)]),
    "vflip_solver_destroy": (c_int, [P]),
    "vflip_solver_set_opts": (c_int, [P, ctypes.POINTER(SolveOpts)]),
    "vflip_solver_set_stream": (c_int, [P, P]),
    "vflip_solver_levels": (c_int, [P]),
    "vflip_solver_level_dims": (c_int, [P, c_int, PI, PI, PI]),
    "vflip_solver_set_matrix": (c_int, [P, P, P, P, P, P]),
    "vflip_solver_setup": (c_int, [P]),
    "vflip_solver_solve": (c_int, [P, P, P, ctypes.POINTER(SolveStats)]),
    "vflip_batch_create": (c_int, [c_int, P, ctypes.POINTER(c_void_p)]),
    "vflip_assemble_liquid": (c_int, [P, c_float, c_float, P, P, P, P, P, P, P, P, P, P, P]),
    "vflip_liquid_phi": (c_int, [c_int, c_int, c_int, c_float, c_float, P, P, P, P, P]),
    "vflip_reorder": (c_int, [P, c_int, c_float, c_float, c_float, ctypes.c_uint, c_int, P, P, P, P, P, P, P, P,
                              c_int, P, P, PI]),
    "vflip_nested_velocity": (c_int, [ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), c_int, P, P, P]),
    "vflip_advect_particles_nested": (c_int, [ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), c_float,
                                              c_int, P, P]),
    "vflip_halo_select": (c_int, [P, ctypes.POINTER(SubdomainView), c_float, c_int, P, P, c_int, P, P, PI]),
    "vflip_seed_ghost": (c_int, [P, ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), ctypes.c_uint, c_int,
                                 P, P, PI]),
    "vflip_pse_diffuse": (c_int, [ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), c_float, c_float, c_int,
                                  P, P, P, P, P, P]),
    "vflip_advect_tracers": (c_int, [ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), c_float, c_float,
                                     c_int, P, P]),
    "vflip_emit_tracers_sphere": (c_int, [P, ctypes.c_uint, c_int, c_float, c_float, c_float, c_float, c_float, P]),
    "vflip_emit_tracers_sub": (c_int, [P, ctypes.POINTER(CoarseView), ctypes.POINTER(SubdomainView), ctypes.c_uint,
                                       c_int, P, PI]),
    "vflip_tracer_frame": (c_int, [c_int, P, c_float, P, P]),
    "vflip_batch_destroy": (c_int, [P]),
    "vflip_batch_set_graph_iterations": (c_int, [P, c_int]),
    "vflip_batch_solve": (c_int, [P, P, P, P, c_int, ctypes.POINTER(ctypes.c_double)]),
    "vflip_batch_autotune": (c_int, [P, P, P, c_int]),
    "vflip_solver_vcycle": (c_int, [P, P, P]),
    "vflip_solver_matvec": (c_int, [P, P, P]),
    "vflip_solver_get_level": (c_int, [P, c_int, P, P, P, P]),
    "vflip_solver_solve_f64": (c_int, [P, P, P, ctypes.POINTER(SolveStats)]),
    "vflip_solver_set_diag64": (c_int, [P, P]),
    "vflip_grad_apply_f64": (c_int, [c_int, c_int, c_int, c_int, c_float, c_float] + [P] * 10 + [P]),
    "vflip_solver_profile": (c_int, [P, c_int]),
    "vflip_solver_profile_read": (c_int, [P, PI, PD, PD]),
    "vflip_amgpcg_solve_host": (c_int, [c_int, c_int, c_int, P, P, P, P, P, P, c_double, c_int, PD, PI,
                                        ctypes.POINTER(SolveOpts)]),
    "vflip_face_weights": (c_int, [c_int, c_int, c_int, P, P, P, P, P]),
    "vflip_inflow_bc": (c_int, [c_int, c_int, c_int, P, c_float, P, P, P, P]),
    "vflip_set_boundary_phi": (c_int, [c_int, c_int, c_int, c_float, P, P, P, P]),
    "vflip_set_boundary_vel": (c_int, [c_int, c_int, c_int, c_float, P, P, P, P, P, P]),
    "vflip_assemble": (c_int, [P, c_int, c_float, c_float] + [P] * 10),
    "vflip_grad_apply": (c_int, [c_int, c_int, c_int, c_int, c_float, c_float] + [P] * 10 + [P]),
    "vflip_project": (c_int, [P, c_int, c_float, c_float] + [P] * 11 + [ctypes.POINTER(SolveStats)]),
    "vflip_advect_rk2": (c_int, [c_int, c_int, c_int, c_float, c_float, c_int, P, P, P, P, c_int, c_int, c_int,
                                 c_float, c_float, c_float, P, P, P, P]),
    "vflip_advect_plume": (c_int, [c_int, c_int, c_int, c_float, c_float] + [P] * 10 + [P]),
    "vflip_plume_force": (c_int, [c_int, c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_float,
                                  ctypes.c_uint, c_float, c_float, c_float, P, P, P, P]),
    "vflip_particles_create": (c_int, [c_int, c_int, c_int, c_float, ctypes.POINTER(c_void_p)]),
    "vflip_particles_destroy": (c_int, [P]),
    "vflip_particles_set_stream": (c_int, [P, P]),
    "vflip_hash": (c_int, [P, c_int, P, c_int, c_float, c_float, c_float, P, P, P, PI]),
    "vflip_hash_noalloc": (c_int, [c_int, c_int, c_int, c_float, c_int, P, c_int, c_float, c_float, c_float, P, P, P,
                                   P]),
    "vflip_p2g": (c_int, [P, c_int, c_int, c_int, P, P, P, P, c_float, c_float, c_float] + [P] * 6),
    "vflip_get_delta": (c_int, [ctypes.c_size_t, P, P, P, P]),
    "vflip_g2p": (c_int, [c_int, c_int, c_int, c_float, c_int, P, P] + [P] * 6 + [c_float] * 4 + [P]),
    "vflip_advect_particles": (c_int, [c_int, c_int, c_int, c_float, c_float, c_int, c_float, c_float, c_float,
                                       c_int, P, P, P, P, P, P]),
    "vflip_fine_to_coarse": (c_int, [P, c_int, c_int, c_int, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, P]),
    "vflip_coarse_to_fine": (c_int, [P, c_int, c_int, c_int, c_float, c_float, c_float, c_float, P, c_int, c_int,
                                     c_int, c_float, c_float, c_float, c_float, P, P]),
    "vflip_set_face_boundary": (c_int, [P, c_int, c_int, c_int, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                        P]),
    "vflip_to_bricks": (c_int, [P, c_int, c_int, c_int, P, P]),
    "vflip_from_bricks": (c_int, [P, c_int, c_int, c_int, P, P]),
    "vflip_unet_create": (c_int, [ctypes.c_uint, ctypes.POINTER(c_void_p)]),
    "vflip_unet_destroy": (c_int, [P]),
    "vflip_unet_set_stream": (c_int, [P, P]),
    "vflip_unet_num_params": (ctypes.c_size_t, [P]),
    "vflip_unet_set_gain": (c_int, [P, c_float]),
    "vflip_unet_gain": (c_float, [P]),
    "vflip_unet_get_params": (c_int, [P, P]),
    "vflip_unet_set_params": (c_int, [P, P]),
    "vflip_unet_get_grads": (c_int, [P, P]),
    "vflip_unet_zero_grad": (c_int, [P]),
    "vflip_unet_get_activation": (c_int, [P, c_int, P]),
    "vflip_unet_forward": (c_int, [P, c_int, c_int, c_int, P, P, P, c_float, P]),
    "vflip_unet_backward": (c_int, [P, P]),
    "vflip_unet_adam": (c_int, [P, c_float, c_int, c_float]),
    "vflip_unet_train_step": (c_int, [P, c_int, c_int, c_int, P, P, P, c_float, P, PD]),
    "vflip_unet_warm_start": (c_int, [P, c_int, c_int, c_int, c_int, P, P, P]),
    "vflip_unet_train_step_energy": (c_int, [P, c_int, c_int, c_int, P, P, P, P, P, P, P, P, PD]),
    "vflip_project_learned": (c_int, [P, P, c_int, c_int, c_float, c_float] + [P] * 11 +
                              [ctypes.POINTER(SolveStats)]),
    "vflip_project_history": (c_int, [P, P, c_int, c_int, c_float, c_float] + [P] * 11 +
                              [ctypes.POINTER(SolveStats)]),
    "vflip_solver_reset_history": (c_int, [P]),
    "vflip_solver_history_guess": (c_int, [P, P, P]),
    "vflip_batch_solve_refined": (c_int, [P, P, P, P, c_int, c_double, c_int, P, c_int, PD]),
    "vflip_solver_get_assembled64": (c_int, [P, P, P]),
    "vflip_solver_solve_refined": (c_int, [P, P, P, P, c_int, c_double, c_int, ctypes.POINTER(SolveStats)]),
    "vflip_mg3d_create": (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "vflip_mg3d_destroy": (c_int, [P]),
    "vflip_mg3d_levels": (c_int, [P, PI, PI]),
    "vflip_mg3d_vcycle": (c_int, [P, P, P, P]),
    "vflip_mg3d_full_multigrid": (c_int, [P, P, P, P]),
    "vflip_mg3d_rbgs": (c_int, [c_int, c_int, c_int, c_int, c_int, P, P, P, c_int, c_double, P]),
    "vflip_mg3d_residual": (c_int, [c_int, c_int, c_int, c_int, c_int, P, P, P, P]),
    "vflip_mg3d_restrict": (c_int, [c_int, c_int, c_int, c_int, P, P, P]),
    "vflip_mg3d_prolongate": (c_int, [c_int, c_int, c_int, c_int, P, P, P]),
}

_lib = None


def lib():
    """Loads libvflip.so; raises if the HIP build is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} not found: build the HIP library first (python -m vflip_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(SIGNATURES)


class VflipError(RuntimeError):
    pass


ERR_CAPACITY = -5


def check(rc, what=""):
    """Raises VflipError for negative status codes; returns rc otherwise."""
    if rc < 0:
        msg = lib().vflip_last_error().decode(errors="replace")
        raise VflipError(f"{what} failed ({rc}): {msg}")
    return rc
