/*
 * vflip.h — C ABI of the MI355X-native vFLIP hot path (libvflip.so).
 *
 * Plain pointers and sizes only. Unless a function says "host", every array
 * argument is a DEVICE pointer (HBM), laid out linear with x fastest:
 * a cell grid ni x nj x nk has index i + ni*(j + nj*k); MAC faces use the
 * staggered shapes u:(ni+1,nj,nk), v:(ni,nj+1,nk), w:(ni,nj,nk+1) and the nodal
 * solid SDF is (ni+1,nj+1,nk+1), all with the reference's world placement
 * (cell (i,j,k) centre at (i,j,k)*h, u(i,j,k) at ((i-.5)h, jh, kh), nodal
 * phi(i,j,k) at ((i-.5)h,(j-.5)h,(k-.5)h)).
 *
 * Status codes: 0 ok/converged, 1 not converged, <0 error (vflip_last_error()).
 * No exceptions cross the ABI. Work is enqueued on the stream set on the
 * object (default: the null stream); calls return after the work completes
 * unless documented otherwise.
 *
 * Reference interfaces replaced (paths relative to the vFLIP checkout,
 * S/ = Smoke_Solver/) are cited on each declaration.
 */
#ifndef VFLIP_H
#define VFLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFLIP_OK 0
#define VFLIP_NOT_CONVERGED 1
#define VFLIP_ERR_ARG (-1)
#define VFLIP_ERR_HIP (-2)
#define VFLIP_ERR_NOMEM (-3)
#define VFLIP_ERR_BREAKDOWN (-4) /* rho == 0 or NaN after the first preconditioner apply */
#define VFLIP_ERR_CAPACITY (-5)  /* an output does not fit the capacity the caller gave */

/* Last error message of the calling thread ("" if none). */
const char* vflip_last_error(void);
/* Library version string. */
const char* vflip_version(void);
/* Number of visible HIP devices (0 when none). */
int vflip_device_count(void);

/* ------------------------------------------------------------------ solver */
typedef struct vflip_solver vflip_solver;

enum vflip_precision {
    VFLIP_PREC_FP32 = 0,  /* plain: FP32 storage, FP32 smoothing, FP32 CG reductions */
    VFLIP_PREC_MIXED = 1  /* FP32 smoothing on FP16-stored levels >= half_from_level, FP64
                             CG dots and alpha/beta, FP64 true-residual refinement every
                             refine_every iterations; stops on the FP64 true residual */
};

enum vflip_variant {
    VFLIP_VARIANT_COARSE = 0, /* SmokeSolver3D::pcg_projection, S/Smoke_solver3D.cpp:945-1200 */
    VFLIP_VARIANT_SUB = 1     /* SubDomainSolver::pcg_projection, S/sub_domain.cpp:556-820 */
};

typedef struct {
    double tolerance;   /* stop when |r|_inf <= tolerance * |b|_inf (recursive r) */
    int max_iterations; /* 50 coarse grid / 1000 subdomain in the reference */
    int precision;      /* enum vflip_precision */
    int pre_sweeps;     /* RB-GS sweeps before restriction (reference: 4) */
    int post_sweeps;    /* RB-GS sweeps after prolongation (reference: 4) */
    int coarse_sweeps;  /* RB-GS sweeps on the coarsest level (reference: 200) */
    int refine_every;   /* mixed: recompute r = b - A x in FP64 every N iterations (4) */
    int warm_start;     /* 1: the x passed to vflip_solver_solve is the initial guess */
    int half_from_level; /* mixed: the V-cycle reads FP16 copies of levels >= this (0: all;
                            the PCG matvec and the true residual keep FP32 level 0) */
} vflip_solve_opts;

typedef struct {
    int iterations;       /* as AMGPCGSolve iterations_out */
    int converged;        /* 1 when the stop test was met */
    double residual;      /* final recursive |r|_inf, absolute (AMGPCGSolve residual_out) */
    double rel_residual;  /* residual / |b|_inf */
    double true_residual; /* |b - A x|_inf / |b|_inf recomputed in FP64 on the GPU */
    double ms_setup;      /* hierarchy build (coarsening kernels) */
    double ms_solve;      /* PCG loop */
    int warm_rejected;    /* warm_start only: 1 when the guess was no better than zero in the
                             A-norm (E(x0) = x0'Ax0/2 - b'x0 >= 0) and the solve restarted from 0 */
    double guess_residual; /* |b - A x0|_inf / |b|_inf of the initial iterate the PCG started
                              from (1 for a zero start; FP64 for a warm start) */
} vflip_solve_stats;

/* Reference defaults: tol 1e-6, 50 iterations, FP32, 4/4/200 sweeps; mixed-mode
 * defaults refine_every 4, half_from_level 0. */
void vflip_default_opts(vflip_solve_opts* opts);

/* Allocates the multigrid hierarchy (levels ceil-halved while > 16^3 unknowns,
 * S/GeometricLevelGen.h:28-71) and the PCG workspace for an ni x nj x nk grid. */
int vflip_solver_create(int ni, int nj, int nk, const vflip_solve_opts* opts, vflip_solver** out);
int vflip_solver_destroy(vflip_solver* s);
int vflip_solver_set_opts(vflip_solver* s, const vflip_solve_opts* opts);
/* hip_stream: a hipStream_t (NULL = null stream). */
int vflip_solver_set_stream(vflip_solver* s, void* hip_stream);
/* Number of levels and the dims of level l (for tests and tooling). */
int vflip_solver_levels(const vflip_solver* s);
int vflip_solver_level_dims(const vflip_solver* s, int level, int* ni, int* nj, int* nk);

/* Level-0 operator from device FP32 arrays: diag and the couplings to the +x,
 * +y, +z neighbours (ax[c] = A(c,c+1) = A(c+1,c); 0 where no entry exists).
 * leak may be NULL; otherwise leak[c] = diag[c] - sum of |couplings| of row c
 * (the Dirichlet part), which keeps coarse diagonals cancellation-free. */
int vflip_solver_set_matrix(vflip_solver* s, const float* diag, const float* ax, const float* ay,
                            const float* az, const float* leak);

/* Builds coarse levels A_{l+1} = 0.5 R A_l P (R = 1/8 average, P = injection),
 * S/GeometricLevelGen.cpp:8-76, in closed form on the GPU. */
int vflip_solver_setup(vflip_solver* s);

/* Preconditioned CG with one V-cycle per iteration (S/AlgebraicMultigrid.h:189-301).
 * rhs, x: device FP32 arrays of ni*nj*nk. Returns 0 converged / 1 not converged /
 * VFLIP_ERR_BREAKDOWN. stats may be NULL. */
int vflip_solver_solve(vflip_solver* s, const float* rhs, float* x, vflip_solve_stats* stats);

/* FP64 defect correction around this solver's own solves: the single-grid form of
 * vflip_batch_solve_refined (same rounds, same kernels; a batch of solvers with the
 * default tile shapes gives the same bits). */
int vflip_solver_solve_refined(vflip_solver* s, const float* rhs, float* x, double* x64, int assembled, double tol,
                               int max_rounds, vflip_solve_stats* stats);
/* The FP64 level-0 diagonal and right-hand side the last assembly wrote (device
 * doubles, ni*nj*nk each; either may be NULL), for checks of FP64 results. */
int vflip_solver_get_assembled64(vflip_solver* s, double* diag, double* rhs);
/* FP64 right-hand side and solution (device double arrays). In mixed precision
 * the iterate is accumulated in FP64 and the stop test is on the FP64 true
 * residual |b - A x|_inf (FP64 diagonal from vflip_solver_set_diag64 or the
 * assembly when available); in FP32 mode the rhs is rounded to FP32 and the
 * FP32 solution is returned widened. Same return codes as vflip_solver_solve. */
int vflip_solver_solve_f64(vflip_solver* s, const double* rhs, double* x, vflip_solve_stats* stats);

/* Exact FP64 level-0 diagonal (device, ni*nj*nk doubles) for the FP64 true
 * residual; the FP32 couplings are exact when they are single FP32 terms, as
 * the reference's assembly produces. Valid until the next set_matrix. */
int vflip_solver_set_diag64(vflip_solver* s, const double* diag);

/* One V-cycle z = M^-1 r (zero initial guess) — the preconditioner alone,
 * S/AlgebraicMultigrid.h:130-188. */
int vflip_solver_vcycle(vflip_solver* s, const float* r, float* z);

/* y = A x with the level-0 operator (S/sparse_matrix.h:341-366). */
int vflip_solver_matvec(vflip_solver* s, const float* x, float* y);

/* Kernel timing of the dominant kernel (the fused level-0 smoothing sweep in
 * its common form: the sweeps after the first of each pre/post group):
 * enable=1 records HIP events around each launch on the solver's stream;
 * _read returns the launches since enabling/the last read, their summed
 * duration and the algorithmic HBM bytes of one launch (28 B per cell; 20 B per
 * cell when a mixed-precision solver reads level 0's FP16 coefficient copy). */
int vflip_solver_profile(vflip_solver* s, int enable);
int vflip_solver_profile_read(vflip_solver* s, int* launches, double* total_ms, double* bytes_per_launch);

/* Copies level l's operator to device arrays (diag, ax, ay, az; FP32). */
int vflip_solver_get_level(const vflip_solver* s, int level, float* diag, float* ax, float* ay,
                           float* az);

/* Drop-in for ivock::AMGPCGSolve<double> (S/AlgebraicMultigrid.h:189-197), HOST
 * arrays: the 7-point SPD matrix as diag + couplings (the reference SparseMatrix
 * rows restricted to a 7-point stencil), rhs and result of ni*nj*nk doubles.
 * Solves on the GPU with the reference stop test; result, residual_out and
 * iterations_out as the reference. opts may be NULL: the mixed-precision solver
 * (VFLIP_PREC_MIXED; FP64 iterate, dots and true-residual stop test, so
 * residual_out is the FP64 |b - A x|_inf) with the default sweeps;
 * tolerance_factor and max_iterations always apply. opts->precision =
 * VFLIP_PREC_FP32 selects plain FP32. Returns 0 converged, 1 not converged,
 * <0 error. */
int vflip_amgpcg_solve_host(int ni, int nj, int nk, const double* diag, const double* ax,
                            const double* ay, const double* az, const double* rhs, double* result,
                            double tolerance_factor, int max_iterations, double* residual_out,
                            int* iterations_out, const vflip_solve_opts* opts);

/* -------------------------------------------------------------- projection */
/* Cut-cell face weights w = clamp(1 - fraction_inside(4 nodal phi), 0, 1);
 * bit-exact with S/Smoke_solver3D.cpp:874-944 / S/levelset_util.cpp:23-103. */
int vflip_face_weights(int ni, int nj, int nk, const float* phi, float* uw, float* vw, float* ww,
                       void* hip_stream);

/* Host callbacks of the boundary set-up: pos points to 3 floats (x, y, z) in
 * world coordinates, user is passed through. They are called from the calling
 * thread, once per grid point. */
typedef float (*vflip_phi_fn)(const float* pos, void* user);
typedef void (*vflip_vel_fn)(const float* pos, float* vel_out, void* user);

/* set_boundary_phi — S/Smoke_solver3D.cpp:851-873: nodal_phi (device,
 * (ni+1)(nj+1)(nk+1)) receives phi(pos) at every node, pos = Vec3f(i,j,k)*h -
 * Vec3f(.5,.5,.5)*h in float. Synchronous. */
int vflip_set_boundary_phi(int ni, int nj, int nk, float h, vflip_phi_fn phi, void* user, float* nodal_phi,
                           void* hip_stream);

/* set_boundary_vel — S/Smoke_solver3D.cpp:793-848: us/vs/ws (device MAC shapes)
 * receive component 0/1/2 of vel(pos) at their faces, pos = Vec3f(i,j,k)*h -
 * Vec3f(ox,oy,oz)*h with the face grids' offsets (.5,0,0), (0,.5,0), (0,0,.5).
 * Synchronous. */
int vflip_set_boundary_vel(int ni, int nj, int nk, float h, vflip_vel_fn vel, void* user, float* us, float* vs,
                           float* ws, void* hip_stream);

/* Inflow boundary: faces of cells with bdesc==1 get u=u_in, v=w=0
 * (S/Smoke_solver3D.cpp:964-980). bdesc: ni*nj*nk bytes. */
int vflip_inflow_bc(int ni, int nj, int nk, const int8_t* bdesc, float u_in, float* u, float* v,
                    float* w, void* hip_stream);

/* Assembles the level-0 operator of the solver's grid and the right-hand side
 * rhs (device FP32, ni*nj*nk) from the MAC velocity, the face weights and the
 * solid velocity (S/Smoke_solver3D.cpp:983-1111; variant SUB: S/sub_domain.cpp:612-729). */
int vflip_assemble(vflip_solver* s, int variant, float h, float dt, const float* u, const float* v,
                   const float* w, const float* uw, const float* vw, const float* ww,
                   const float* us, const float* vs, const float* ws, float* rhs);

/* vflip_assemble (coarse variant) with a free surface: cells where liquid_phi >= 0
 * are air, a Dirichlet p = 0 boundary like the domain ring (empty rows), and
 * liquid cells couple only to liquid neighbours (first-order free surface; no
 * reference counterpart, the reference's smoke assembly is the all-liquid case).
 * liquid_phi: device ni*nj*nk (vflip_liquid_phi). */
int vflip_assemble_liquid(vflip_solver* s, float h, float dt, const float* u, const float* v, const float* w,
                          const float* uw, const float* vw, const float* ww, const float* us, const float* vs,
                          const float* ws, const float* liquid_phi, float* rhs);

/* Velocity update u -= grad p (S/Smoke_solver3D.cpp:1126-1200; variant SUB with
 * dt and wider face ranges: S/sub_domain.cpp:745-819). p: ni*nj*nk FP32. */
int vflip_grad_apply(int ni, int nj, int nk, int variant, float h, float dt, const float* p,
                     const float* uw, const float* vw, const float* ww, const float* us,
                     const float* vs, const float* ws, float* u, float* v, float* w,
                     void* hip_stream);

/* Same update from an FP64 pressure: (float)(p[c] - p[c-1]) as the reference,
 * whose pressure vector is double. */
int vflip_grad_apply_f64(int ni, int nj, int nk, int variant, float h, float dt, const double* p,
                         const float* uw, const float* vw, const float* ww, const float* us,
                         const float* vs, const float* ws, float* u, float* v, float* w,
                         void* hip_stream);

/* Whole pressure projection — the body of SmokeSolver3D::pcg_projection /
 * SubDomainSolver::pcg_projection after face weights: assemble, setup, PCG,
 * gradient update. pressure (device, ni*nj*nk) receives p; it is also the warm
 * start when opts.warm_start is set. rhs_scratch: device ni*nj*nk FP32. */
int vflip_project(vflip_solver* s, int variant, float h, float dt, float* u, float* v, float* w,
                  const float* uw, const float* vw, const float* ww, const float* us,
                  const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                  vflip_solve_stats* stats);

/* --------------------------------------------------------------- advection */
/* Semi-Lagrangian advection of one field (and optionally a second field on the
 * same grid, sharing the backtrace) by the MAC velocity u,v,w of the ni x nj x nk
 * cell grid: RK2 midpoint backtrace x - dt*u(x - dt/2*u(x)), then
 *   mode 0: trilinear gather — SubDomainSolver::advect_field, S/sub_domain.cpp:191-235
 *   mode 1: cubic gather clamped to the trilinear value on overshoot —
 *           SmokeSolver3D::advect_field_cubic_clamp, S/Smoke_solver3D.cpp:657-707.
 * field: fnx x fny x fnz with grid offsets (fox,foy,foz) in cells (u: .5,0,0;
 * cell-centred scalars: 0,0,0). Samplers interpolate in FP64 with index clamping
 * (Buffer3D::sample_linear/sample_cubic, S/fluid_buffer3D.h:226-326); results are
 * bit-identical with that arithmetic.
 *   mode 2: mode 0's trilinear backtrace and gather evaluated in FP32 (fused
 *           lerps), within a few FP32 ulps of mode 0 and several times faster;
 *           the simulation step's choice.
 * field2/out2 may be NULL. Asynchronous on hip_stream. */
int vflip_advect_rk2(int ni, int nj, int nk, float h, float dt, int mode, const float* u, const float* v,
                     const float* w, const float* field, int fnx, int fny, int fnz, float fox, float foy,
                     float foz, float* out, const float* field2, float* out2, void* hip_stream);

/* The smoke step's five advections in one launch: u, v, w (MAC faces) and the
 * cell-centred rho and T, each by vflip_advect_rk2 mode 2 with the MAC velocity
 * u, v, w (S/sub_domain.cpp:191-235 in FP32; the coarse grid's cubic form is
 * mode 1 of vflip_advect_rk2). Bit-identical with the five mode-2 calls, in one
 * launch (each thread the five outputs of a pair of cells). Outputs must not alias
 * inputs. Asynchronous on hip_stream. */
int vflip_advect_plume(int ni, int nj, int nk, float h, float dt, const float* u, const float* v, const float* w,
                       const float* rho, const float* temp, float* u2, float* v2, float* w2, float* rho2, float* temp2,
                       void* hip_stream);

/* Smoke-plume forcing of the synthetic workload (replaces the reference's
 * gen_heat/add_smokeforce, S/Smoke_solver3D.cpp:1624-1668, with the workload's
 * sphere source and face buoyancy): cells with |x - c| <= r get
 * rho = T = 1 + jitter*U(seed, cell); then v faces 1..nj-1 get
 * v += dt*(-alpha*rho_f + beta*T_f). Asynchronous on hip_stream. */
int vflip_plume_force(int ni, int nj, int nk, float h, float dt, float cx, float cy, float cz, float r,
                      unsigned seed, float jitter, float alpha, float beta, float* rho, float* temp, float* v,
                      void* hip_stream);

/* -------------------------------------------------------------------- FLIP */
/* Particles are xyz-interleaved float arrays pos[3*np], vel[3*np]: the pos and
 * vel members of the reference's particle (S/fluid_particle.h:14-65). The
 * FLIP calls below are asynchronous on their stream unless stated. */
typedef struct vflip_particles vflip_particles;

/* Binning and transfer workspace for an ni x nj x nk cell grid of spacing h
 * (the reference's per-cell particle_hash lists, S/Smoke_solver3D.h:205-212).
 * Device scratch grows with the particle count. */
int vflip_particles_create(int ni, int nj, int nk, float h, vflip_particles** out);
int vflip_particles_destroy(vflip_particles* P);
int vflip_particles_set_stream(vflip_particles* P, void* hip_stream);

/* hash_FLIP — SmokeSolver3D, S/Smoke_solver3D.cpp:1327-1346 (clamp = 0:
 * particles outside the grid are dropped) and SubDomainSolver,
 * S/sub_domain.cpp:1087-1107 (clamp = 1: positions relative to bmin, cell
 * indices clamped into the grid). cell = trunc((p [- bmin] + h/2) / h) in float.
 * Outputs: cell_key[np] (linear cell index, -1 when dropped), cell_start[ncell+1]
 * (offsets into order), order[np] (particle ids grouped by cell, ascending id
 * within a cell = the append order of the reference's per-cell lists).
 * Bit-exact. *nbinned (host, may be NULL) receives cell_start[ncell]; passing it
 * makes the call synchronous. */
int vflip_hash(vflip_particles* P, int np, const float* pos, int clamp, float bx, float by, float bz,
               int* cell_key, int* cell_start, int* order, int* nbinned);

/* vflip_hash without a workspace object and without any allocation. No counterpart
 * in the reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and bit-exact results as vflip_hash on an ni x nj x nk grid of
 * spacing h; the binning's scratch (slots, merge buffer, scan partials) is one
 * static device array per GPU (about 34 MB, held until the process ends). Limits:
 * np <= 4194304 particles and ni*nj*nk <= 16777216 cells; larger inputs return
 * VFLIP_ERR_ARG before any launch. Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams or threads) race. Asynchronous on
 * hip_stream. */
int vflip_hash_noalloc(int ni, int nj, int nk, float h, int np, const float* pos, int clamp, float bx, float by,
                       float bz, int* cell_key, int* cell_start, int* order, void* hip_stream);

/* particle_to_grid — S/Smoke_solver3D.cpp:1451-1510 (VFLIP_VARIANT_COARSE) and
 * S/sub_domain.cpp:1204-1265 (VFLIP_VARIANT_SUB: positions relative to bmin;
 * faces within 4 cells of the boundary use a 4h kernel over 9^3 cells,
 * S/sub_domain.h:89-103). For every MAC face: coef = 1e-5 + sum w,
 * field = sum w*vel_c / coef, w = tent weights H(dx/h)H(dy/h)H(dz/h)
 * (S/Smoke_solver3D.h:217-233) over the particles of the 3^3 cells around the
 * face. Takes vflip_hash's cell_start/order. method 0: owner-computes tiles
 * (per-cell register sums, LDS-binned adds, plain stores); method 1: per-face
 * gather in the reference's summation order (bit-exact); method 2: LDS-binned
 * scatter over each tile's own cells, then global atomics for every face.
 * Methods 0 and 2 differ from the reference at FP32 rounding level only: they
 * sum in another order, and method 0 forms the tent arguments from one scaled
 * offset per axis. A component with NULL field and coef is skipped. */
int vflip_p2g(vflip_particles* P, int variant, int method, int np, const float* pos, const float* vel,
              const int* cell_start, const int* order, float bx, float by, float bz, float* u, float* u_coef,
              float* v, float* v_coef, float* w, float* w_coef);

/* getDelta — S/Smoke_solver3D.cpp:1297-1318: out = fnew - fold (n floats). */
int vflip_get_delta(size_t n, const float* fnew, const float* fold, float* out, void* hip_stream);

/* particle_from_grid — S/Smoke_solver3D.cpp:1398-1406: vel += trilinear(u_new -
 * u_old) at pos - (bx,by,bz), per component (the reference samples _utemp after
 * getDelta; the difference is formed per corner here, with the same float
 * rounding). With u_old == NULL, u_new/v_new/w_new are the delta grids.
 * flip_ratio < 1 blends with PIC in double as S/fluid_particle.h:297-336:
 * vel = (1 - r)*u_new(pos) + r*(vel + du(pos)); it needs the old grids.
 * Bit-exact. */
int vflip_g2p(int ni, int nj, int nk, float h, int np, const float* pos, float* vel, const float* u_new,
              const float* v_new, const float* w_new, const float* u_old, const float* v_old,
              const float* w_old, float flip_ratio, float bx, float by, float bz, void* hip_stream);

/* advect_FLIP — traceRK3 then solid push-out (pos -= phi * normalize(grad phi)
 * where phi < 0). VFLIP_VARIANT_COARSE: S/Smoke_solver3D.cpp:141-154, 1408-1430
 * (c3 = (1 - c1 - c2)*dt with c1 = 2/9 dt, c2 = 3/9 dt already scaled by dt);
 * VFLIP_VARIANT_SUB: S/sub_domain.cpp:116-129, 1180-1202 (c3 = 4/9 dt, velocity
 * at pos - bmin, phi at the unshifted pos). Trilinear velocity (without the
 * subdomain blend of S/Smoke_solver3D.cpp:71-88); solid_phi is the nodal SDF
 * (ni+1, nj+1, nk+1). Bit-exact. */
int vflip_advect_particles(int ni, int nj, int nk, float h, float dt, int variant, float bx, float by, float bz,
                           int np, float* pos, const float* u, const float* v, const float* w,
                           const float* solid_phi, void* hip_stream);

/* reorder_FLIP — SmokeSolver3D, S/Smoke_solver3D.cpp:1347-1396
 * (VFLIP_VARIANT_COARSE) and SubDomainSolver, S/sub_domain.cpp:1110-1159
 * (VFLIP_VARIANT_SUB: interior cells 1..n-2 only, so particles binned in the
 * boundary ring are dropped). Takes vflip_hash's lists (clamp 0 / clamp 1 with
 * the same bmin) and rebuilds the particle set in linear cell order: a cell with
 * n particles keeps them all when n < 10, else its first 8, in list order; when
 * n < 8 it also gets 8 - n new particles at [bmin +] (i,j,k) h + 0.49 h r,
 * r = frand(-1, 1) per axis, with velocity trilinear in u, v, w at that position
 * (get_velocity_self; [minus bmin] in the subdomain form). Then every particle
 * whose solid_phi (nodal SDF, (ni+1)(nj+1)(nk+1), trilinear at pos [- bmin]) is
 * < 0 is dropped. The reference draws r with rand(); here
 * r = -1 + 2 (randhash(seed * 0x9E3779B1 + 3 (8 cell + slot) + axis) >> 8) / 2^24,
 * deterministic per (seed, cell, slot). out_pos/out_vel (3*cap floats, not
 * aliasing pos/vel) receive the new set, *nout (host) its size. When the size
 * exceeds cap nothing is written and VFLIP_ERR_CAPACITY is returned with *nout
 * set; np + 8 * (cells) always suffices. Synchronous on the workspace's stream.
 * Bit-exact. */
int vflip_reorder(vflip_particles* P, int variant, float bx, float by, float bz, unsigned seed, int np,
                  const float* pos, const float* vel, const int* cell_start, const int* order, const float* u,
                  const float* v, const float* w, const float* solid_phi, int cap, float* out_pos, float* out_vel,
                  int* nout);

/* Particle level set of the liquid for the free-surface pressure boundary (the
 * reference has no liquid solver): phi at every cell centre (i h, j h, k h) =
 * min |centre - p| over the particles in the 3x3x3 cells around it (vflip_hash's
 * cell_start/order, clamp 0), minus radius; 3h - radius where none is that close.
 * Liquid cells are phi < 0. Exact (float distances). Asynchronous. */
int vflip_liquid_phi(int ni, int nj, int nk, float h, float radius, const float* pos, const int* cell_start,
                     const int* order, float* phi, void* hip_stream);

/* ----------------------------------------------- nested FLIP coupling */
/* The coarse solver's grids (SmokeSolver3D): ni x nj x nk cells of spacing h,
 * MAC velocity u (ni+1,nj,nk), v (ni,nj+1,nk), w (ni,nj,nk+1), nodal solid phi
 * ((ni+1)(nj+1)(nk+1)), MAC solid velocity us, vs, ws (NULL = zero). Device
 * pointers, linear x-fastest. */
typedef struct vflip_coarse_view {
    int ni, nj, nk;
    float h;
    const float *u, *v, *w, *solid_phi, *us, *vs, *ws;
} vflip_coarse_view;

/* One subdomain as the coarse solver sees it (SubDomainSolver members,
 * S/sub_domain.h:160-170): cells, spacing h (_hx), bmin and bmax as the
 * reference stores them, and the same fields on the subdomain grid, which its
 * samplers read at pos - bmin (S/sub_domain.cpp:60-95). */
typedef struct vflip_subdomain_view {
    int ni, nj, nk;
    float h;
    float bmin[3], bmax[3];
    const float *u, *v, *w, *solid_phi, *us, *vs, *ws;
} vflip_subdomain_view;

/* get_velocity with the subdomain blend — S/Smoke_solver3D.cpp:71-88,
 * S/sub_domain.cpp:70-77 and 96-105: the coarse trilinear velocity; inside
 * is_in2 (bmin + 5h < pos < bmax - 5h, in double; S/sub_domain.h:190-201)
 * c * v_sub + (1 - c) * v with c = clamp(d / 4h, 0, 1), d the least distance to
 * the faces of that box. out[3*n]. Asynchronous. Bit-exact. */
int vflip_nested_velocity(const vflip_coarse_view* C, const vflip_subdomain_view* S, int n, const float* pos,
                          float* out, void* hip_stream);

/* advect_FLIP of the coarse particles with a subdomain — S/Smoke_solver3D.cpp:
 * 141-154 (traceRK3, c3 = (1 - c1 - c2) dt) through the blended velocity, and
 * 1408-1430 (pos -= phi * normal where phi < 0) with get_solid_phi /
 * get_solid_normal taking the subdomain's phi inside is_in (margin 2h,
 * S/Smoke_solver3D.cpp:106-124). pos updated in place. Asynchronous. Bit-exact. */
int vflip_advect_particles_nested(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, int np,
                                  float* pos, void* hip_stream);

/* The halo copy of S/Smoke_solver3D.cpp:393-401: the particles with
 * is_halo(pos, dx) (bmin - dx < pos < bmax + dx in float, S/sub_domain.h:203-214)
 * in their order. P: any particle workspace (its scratch and stream).
 * Synchronous; capacity as vflip_reorder (np always fits). Bit-exact. */
int vflip_halo_select(vflip_particles* P, const vflip_subdomain_view* S, float dx, int np, const float* pos,
                      const float* vel, int cap, float* out_pos, float* out_vel, int* nout);

/* seed_ghost — S/Smoke_solver3D.cpp:1529-1560: for every subdomain cell whose
 * get_solid_phi at bmin + (i,j,k) h is < 3h, 8 candidates at
 * (i,j,k) h + 0.49 h r + bmin; a candidate whose subdomain phi lies in
 * (-2h, 2h] is kept, with the blended get_velocity there, or get_solid_vel where
 * that phi is <= 0. Cell-major order. r = frand(-1, 1) per axis drawn as in
 * vflip_reorder with key 3 (8 cell + slot) + axis (the reference uses rand()).
 * P: the subdomain's workspace (vflip_particles_create over its cells).
 * Synchronous; capacity as vflip_reorder (8 * cells always fits). Bit-exact. */
int vflip_seed_ghost(vflip_particles* P, const vflip_coarse_view* C, const vflip_subdomain_view* S, unsigned seed,
                     int cap, float* out_pos, float* out_vel, int* nout);

/* diffuse (PSE viscosity) — S/Smoke_solver3D.cpp:1562-1621 with PSE_kernel of
 * S/Smoke_solver3D.h:204-209: for each of the first np_real subdomain particles
 * (pos/vel continue with the ghosts of vflip_seed_ghost), when get_solid_phi at
 * it is <= 2 h_coarse and its cell trunc((x - bmin)/h + 0.5) lies in the
 * subdomain, du = sum_j (v_j - v_i) K_ij / (1e-9 + sum_j K_ij) over the particles
 * that vflip_hash (subdomain workspace, clamp 1, bmin) bins in the 3x3x3 cells
 * around it, in the reference's order; tau = dt nu + 1e-10. out_vel[3*np_real]
 * = vel + du (not aliasing vel); every real particle must be in the bins, as
 * the clamped binning puts every particle. Asynchronous. The kernel's exponential is the
 * device's double exp: within one float rounding of the host's per term. */
int vflip_pse_diffuse(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, float nu, int np_real,
                      const float* pos, const float* vel, const int* cell_start, const int* order, float* out_vel,
                      void* hip_stream);

/* ------------------------------------------------------------- tracers */
/* Smoke tracers are float4 (x, y, z, density) arrays, 16-byte aligned: the
 * reference's Vec4f tracers (S/Smoke_solver3D.h:45-161). */

/* advect_tracers — S/Smoke_solver3D.h:101-125: traceRK3 through get_velocity
 * (with the subdomain blend when S is not NULL; coarse c3 = (1 - c1 - c2) dt),
 * pos -= phi * normal where get_solid_phi <= 0, density d / (1 + decay dt) in
 * double. In place. Asynchronous. Bit-exact. */
int vflip_advect_tracers(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, float dens_decay,
                         int n, float* tracers, void* hip_stream);

/* The sphere emitter of emit_tracers — S/Smoke_solver3D.h:67-82: candidates
 * (x, y, z) = frand(-r - hx, r + hx) per axis, accepted when
 * x^2 + z^2 + y^2 <= r^2; the first n_emit accepted in candidate order become
 * tracers (cx + x, cy + y, cz + z, 1). frand(a, b) = a + (b - a) U with
 * U = (randhash(seed * 0x9E3779B1 + 3 t + axis) >> 8) / 2^24 for candidate t
 * (the reference draws rand()). out[4 * n_emit]. P: any particle workspace
 * (scratch and stream). Synchronous. Bit-exact. */
int vflip_emit_tracers_sphere(vflip_particles* P, unsigned seed, int n_emit, float cx, float cy, float cz, float r,
                              float hx, float* out);

/* The subdomain part of emit_tracers — S/Smoke_solver3D.h:84-96: per subdomain
 * cell (cell-major) 4 candidates bmin + (i,j,k) h + U h (U = frand(0, 1) per
 * axis, key 3 (4 cell + slot) + axis), kept where 0 < get_solid_phi < h, as
 * tracers with density 1. P: the subdomain's workspace. Synchronous; capacity
 * as vflip_reorder (4 * cells always fits). Bit-exact. */
int vflip_emit_tracers_sub(vflip_particles* P, const vflip_coarse_view* C, const vflip_subdomain_view* S,
                           unsigned seed, int cap, float* out, int* nout);

/* The record of write_tracers — S/Smoke_solver3D.h:126-148: out = (x/lx - 0.5,
 * y/lx, z/lx - 0.5, d) per tracer; the file Particle_data%04d.bin is these
 * records, raw float32. Asynchronous. Bit-exact. */
int vflip_tracer_frame(int n, const float* tracers, float lx, float* out, void* hip_stream);

/* ------------------------------------------------- coarse/fine coupling */
/* FineToCoarseOperator — S/coarse_fine_solver.cpp:73-108: coarse(i,j,k) = mean of
 * the fine cells [i*ratio - offx, i*ratio + ratio - offx) x ... clipped to the
 * fine grid (0 if none); FineToCoarse (110-118) uses off = ratio/2 along a MAC
 * face's normal. Bit-exact. Asynchronous on hip_stream. */
int vflip_fine_to_coarse(const float* fine, int fnx, int fny, int fnz, float* coarse, int cnx, int cny, int cnz,
                         int ratio, int offx, int offy, int offz, void* hip_stream);

/* CoarseToFineOperator — S/coarse_fine_solver.cpp:146-170: fine(i,j,k) =
 * coarse.sample_linear(((i,j,k) - (fox,foy,foz))*hf) with the coarse grid's
 * offsets and spacing hc. With fine_minus != NULL it writes
 * fine_minus - sample instead (ComputeSmallDifference, 47-71). Bit-exact.
 * Asynchronous on hip_stream. */
int vflip_coarse_to_fine(float* fine, int fnx, int fny, int fnz, float fox, float foy, float foz, float hf,
                         const float* coarse, int cnx, int cny, int cnz, float cox, float coy, float coz, float hc,
                         const float* fine_minus, void* hip_stream);

/* SetFaceBoundary — S/coarse_fine_solver.cpp:182-230: for each coarse cell, the
 * fine samples [i*ratio, i*ratio + offx) x [j*ratio, j*ratio + offy) x ...
 * (the fine faces under one coarse face: extent 1 along its normal) get
 * coarse - sum/(ratio*ratio) added. Requires 0 <= off <= ratio. Bit-exact.
 * Asynchronous on hip_stream. */
int vflip_set_face_boundary(float* fine, int fnx, int fny, int fnz, const float* coarse, int cnx, int cny, int cnz,
                            int ratio, int offx, int offy, int offz, void* hip_stream);

/* Buffer3D storage — S/fluid_buffer3D.h:186-202: 8^3 bricks, element
 * ((K*bx*by + J*bx + I) << 9) + (kk << 6) + (jj << 3) + ii. to_bricks writes
 * bx*by*bz*512 floats (padding 0); from_bricks reads them back to the linear
 * nx*ny*nz layout of this ABI. Asynchronous on hip_stream. */
int vflip_to_bricks(const float* src, int nx, int ny, int nz, float* dst, void* hip_stream);
int vflip_from_bricks(const float* src, int nx, int ny, int nz, float* dst, void* hip_stream);

/* ------------------------------------------- batched subdomain solves */
/* The nested fine subdomains' pressure solves (SubDomainSolver::pcg_projection's
 * AMGPCGSolve call, S/sub_domain.cpp:739, once per subdomain in the reference)
 * run together: every V-cycle and PCG step is one launch over all subdomains,
 * the CG scalars and stop tests stay on the GPU, and a fixed number of PCG
 * iterations is captured once into a hipGraph and replayed until every
 * subdomain has converged. The batch references (does not own) n solvers, each
 * FP32 (VFLIP_PREC_FP32) with its operator set (vflip_solver_set_matrix or
 * vflip_assemble); their per-subdomain tolerance and iteration limit apply. With
 * the solvers' default tile shapes the results equal n vflip_solver_solve calls
 * bit for bit. Rebuild the batch after a solver's operator hierarchy is rebuilt
 * with different options. */
typedef struct vflip_batch vflip_batch;
int vflip_batch_create(int n, vflip_solver* const* solvers, vflip_batch** out);
int vflip_batch_destroy(vflip_batch* batch);
/* PCG iterations per captured graph (1..64, default 8): one host check per replay. */
int vflip_batch_set_graph_iterations(vflip_batch* batch, int iterations);
/* Solves A_q x_q = rhs_q for every subdomain q (device FP32 arrays). use_graph=0
 * issues the same launches directly. stats: n entries or NULL; ms: wall time of
 * the batched PCG (NULL to skip). Returns 0 when all converged, 1 when one did
 * not, <0 on error (VFLIP_ERR_BREAKDOWN if a subdomain broke down). */
int vflip_batch_solve(vflip_batch* batch, const float* const* rhs, float* const* x, vflip_solve_stats* stats,
                      int use_graph, double* ms);
/* The subdomain solves to the reference's FP64 tolerance (SubDomainSolver::pcg_projection
 * solves in FP64 to a relative 1e-9 in at most 1000 iterations, S/sub_domain.cpp:739):
 * FP64 defect correction around the batched FP32 MGPCG. Each round takes every
 * unfinished subdomain's FP64 true residual b - A x (the assembled FP64 diagonal; with
 * `assembled`, the assembled FP64 right-hand side, else rhs), solves all corrections
 * in one vflip_batch_solve to the solvers' own (FP32) tolerance and adds them to the
 * FP64 iterates, until |b - A x|_inf <= tol |b|_inf or max_rounds corrections.
 * x: FP32 solutions; x64 (may be NULL, else per subdomain NULL or device doubles):
 * the FP64 iterates. stats: iterations summed over the rounds, true_residual the
 * final FP64 |b - A x|_inf / |b|_inf, ms_solve the call's wall time. */
int vflip_batch_solve_refined(vflip_batch* batch, const float* const* rhs, float* const* x, double* const* x64,
                              int assembled, double tol, int max_rounds, vflip_solve_stats* stats, int use_graph,
                              double* ms);
/* Autotuner: times the fused smoothing sweep over the whole batch for each level
 * depth and tile shape (4/8/16 rows per tile x z chunks of 4..64 planes), keeps
 * the fastest and rebuilds the batch's launch plan. ty/zc (may be NULL) receive
 * the chosen shape of the first max_levels depths (0 where no level is fused).
 * Tuned shapes change the grouping of the z.r partial sums, so results then
 * agree with vflip_solver_solve to rounding, not bit for bit. */
int vflip_batch_autotune(vflip_batch* batch, int* ty, int* zc, int max_levels);

/* ------------------------------------------------------ learned warm start */
/* No counterpart in the reference: a 3D U-Net (4 levels, 16/32/64/128 channels,
 * 3x3x3 convolutions, BF16 weights/activations, FP32 accumulation on the matrix
 * cores) that maps (rhs * in_scale, solid mask, liquid mask) to a pressure
 * guess for vflip_solver_solve's warm_start; MGPCG still runs to the requested
 * tolerance. Grids: ni, nj, nk positive multiples of 8, cell arrays linear
 * x-fastest (device). Parameters are fp32 [weights of every layer][biases]
 * (vflip_unet_num_params floats, layout in vflip_amd/unet.py). */
typedef struct vflip_unet vflip_unet;
int vflip_unet_create(unsigned seed, vflip_unet** out);
int vflip_unet_destroy(vflip_unet* net);
int vflip_unet_set_stream(vflip_unet* net, void* hip_stream);
size_t vflip_unet_num_params(const vflip_unet* net);
/* Pressure unit of the automatic normalisation (below): pressure / (max|rhs| * gain)
 * is what the network predicts. Set from the training data; default 1. */
int vflip_unet_set_gain(vflip_unet* net, float gain);
float vflip_unet_gain(const vflip_unet* net);
int vflip_unet_get_params(const vflip_unet* net, float* dst);
int vflip_unet_set_params(vflip_unet* net, const float* src);
int vflip_unet_get_grads(const vflip_unet* net, float* dst);
int vflip_unet_zero_grad(vflip_unet* net);
/* A stored activation of the last forward as fp32, [z][y][x][C] on its level's grid
 * (C = the tensor's channels): tensor 0..13 = output of layer 0..13 (after the ReLU),
 * 15..17 = the pooled inputs of levels 1..3, 18 = the packed network input (16
 * channels, 0..2 used). For tests that check the backward pass against a reference
 * fed with the same forward values. Synchronous. */
int vflip_unet_get_activation(const vflip_unet* net, int tensor, float* dst);
/* pred[v] = network output; solid/liquid may be NULL (0 / 1). Asynchronous. */
int vflip_unet_forward(vflip_unet* net, int ni, int nj, int nk, const float* rhs, const float* solid,
                       const float* liquid, float in_scale, float* pred);
/* Adds d(loss)/d(params) for grad_pred = d(loss)/d(pred) of the last forward. */
int vflip_unet_backward(vflip_unet* net, const float* grad_pred);
/* Adam (beta 0.9/0.999, eps 1e-8) on the fp32 master copy with the gradients
 * scaled by grad_scale; step counts from 1. */
int vflip_unet_adam(vflip_unet* net, float lr, int step, float grad_scale);
/* Forward, loss = 0.5 * mean((pred - target)^2), backward (gradients add up);
 * *loss (host, may be NULL; non-NULL synchronises). in_scale > 0: input
 * rhs * in_scale, raw output; in_scale == 0: automatic normalisation by
 * s = max|rhs| (computed on the device), the target taken as
 * pressure / (s * gain) and vflip_unet_forward's output returned as pressure
 * (times s * gain). */
int vflip_unet_train_step(vflip_unet* net, int ni, int nj, int nk, const float* rhs, const float* solid,
                          const float* liquid, float in_scale, const float* target, double* loss);

/* Energy-norm training step (automatic normalisation): loss =
 * |pred - x|_A^2 / |x|_A^2 with x = target / (max|rhs| * gain) and A the sample's
 * 7-point operator (diag, couplings to +x/+y/+z as vflip_solver_set_matrix):
 * the error norm that governs the warm-started PCG. Gradients add up. */
int vflip_unet_train_step_energy(vflip_unet* net, int ni, int nj, int nk, const float* rhs, const float* solid,
                                 const float* liquid, const float* target, const float* diag, const float* ax,
                                 const float* ay, const float* az, double* loss);

/* Pressure guess for a solve on an ni x nj x nk grid: rhs averaged by `factor`
 * per axis onto the network grid (solid fraction from diag == 0), the network
 * in automatic normalisation, the result prolonged trilinearly into x0 (0 on
 * cells with diag == 0). Asynchronous on the network's stream. */
int vflip_unet_warm_start(vflip_unet* net, int ni, int nj, int nk, int factor, const float* rhs, const float* diag,
                          float* x0);

/* vflip_project with the network's guess as the PCG's initial pressure
 * (warm_start for this solve); MGPCG runs to the solver's tolerance as
 * always. factor: solve grid / network grid per axis. */
int vflip_project_learned(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u,
                          float* v, float* w, const float* uw, const float* vw, const float* ww, const float* us,
                          const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                          vflip_solve_stats* stats);

/* vflip_project warm-started from the solver's own pressure history (no reference
 * counterpart: the reference zeroes the pressure before every solve,
 * S/Smoke_solver3D.cpp:960 and S/AlgebraicMultigrid.h:212). The solver keeps the FP64
 * pressures of its last three solves; the initial guess is 0 before the first,
 * else the combination of them that minimises x'Ax/2 - b'x (the Galerkin
 * projection onto previous solutions, Fischer 1998), formed on the GPU. With a
 * network, the divergence that guess leaves, r = b - A x0, goes through the
 * network on the grid coarsened by `factor` (vflip_unet_warm_start with r in
 * place of the right-hand side) and the prolonged correction is added to x0. The
 * PCG vets the guess (dropped when it is no better than zero in the A-norm) and
 * runs to the solver's tolerance as always; the solution becomes the newest
 * history entry (a solve that fails clears the history). net may be NULL: the
 * extrapolation alone. stats->guess_residual reports |b - A x0|_inf / |b|_inf. */
int vflip_project_history(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u,
                          float* v, float* w, const float* uw, const float* vw, const float* ww, const float* us,
                          const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                          vflip_solve_stats* stats);
/* The initial guess vflip_project_history would start from for the operator and
 * right-hand side rhs (device FP32, ni*nj*nk) the solver holds now (its FP64 form
 * when the assembly wrote it), into guess (device FP64, ni*nj*nk). Returns how
 * many history entries it rests on (0: zero guess), <0 on error. Synchronous. */
int vflip_solver_history_guess(vflip_solver* S, const float* rhs, double* guess);
/* Forgets the pressure history (a new scene or a jump in time). */
int vflip_solver_reset_history(vflip_solver* S);

/* ---- Geometric multigrid on a uniform grid: MultiGridSolver3D<T> -----------
 * S/Multigrid3D.h:152-184 (class), 1438-1439 (MGSolverf / MGSolverd). The
 * constant-coefficient 7-point operator (A x)(c) = sum of the six neighbours -
 * 6 x(c), a neighbour past the grid clamped to the cell itself (Neumann) or 0
 * (dirichlet != 0, the reference's _b_Dirichlet). Arrays are device arrays of
 * the chosen type, linear x-fastest (the reference's Buffer3D bricks convert
 * with vflip_to_bricks / vflip_from_bricks for float). Results are
 * bit-identical with the reference's arithmetic in that type. Calls enqueue on
 * `stream` and return after the work completes. */
typedef enum { VFLIP_MG3D_FLOAT = 0, VFLIP_MG3D_DOUBLE = 1 } vflip_mg3d_type;
typedef struct vflip_mg3d vflip_mg3d;
/* Constructor + m_InitialSystem (S/Multigrid3D.h:139-150, levels 100-137): every
 * extent halves (floor, at least 1) until the grid has at most 512 cells; the
 * level buffers (x, b of the coarse levels, x_new of every level) are allocated
 * here. One object serves one stream at a time. */
int vflip_mg3d_create(int gridx, int gridy, int gridz, int type, int dirichlet, vflip_mg3d** out);
int vflip_mg3d_destroy(vflip_mg3d* m);
/* m_max_level and systemk: the level count and, when dims is not NULL, the
 * extents of every level (3 ints per level, finest first). */
int vflip_mg3d_levels(const vflip_mg3d* m, int* levels, int* dims);
/* m_Vcycle(x, b, tol, residual, 0) (S/Multigrid3D.h:1328-1392): one V-cycle on
 * x in place (5 + 5 red-black sweeps per level, 100 at the bottom). The
 * reference's tol and residual are unused there and have no counterpart. */
int vflip_mg3d_vcycle(vflip_mg3d* m, void* x, const void* b, void* stream);
/* m_FullMultiGrid(x, b, tol, residual) (S/Multigrid3D.h:1394-1436). */
int vflip_mg3d_full_multigrid(vflip_mg3d* m, void* x, const void* b, void* stream);
/* m_Red_Black_Gauss(system, b, x, x_new, iter_time, omega) (S/Multigrid3D.h:840-1129)
 * on an nx x ny x nz grid: iter_time sweeps (parity-1 cells, then parity-0) on
 * x in place; x_new is scratch of the same size (its contents are overwritten). */
int vflip_mg3d_rbgs(int nx, int ny, int nz, int type, int dirichlet, const void* b, void* x, void* x_new,
                    int iter_time, double omega, void* stream);
/* m_ComputeResidual(system, res, b, x) (S/Multigrid3D.h:1130-1186): res = b - A x. */
int vflip_mg3d_residual(int nx, int ny, int nz, int type, int dirichlet, void* res, const void* b, const void* x,
                        void* stream);
/* m_Restrict (S/Multigrid3D.h:1193-1245): dst (the coarser grid, extents
 * max(n/2, 1)) = 0.5 * the sum of the 8 children in src (nx x ny x nz). */
int vflip_mg3d_restrict(int nx, int ny, int nz, int type, const void* src, void* dst, void* stream);
/* m_Prolongate (S/Multigrid3D.h:1298-1326): dst (nx x ny x nz) += its parent in
 * src (the coarser grid). A parent past the coarser grid (odd extents) adds 0;
 * the reference reads Buffer3D padding there, which is 0 unless the coarser
 * extent is a multiple of 8 (it then reads a cell of the next brick). */
int vflip_mg3d_prolongate(int nx, int ny, int nz, int type, const void* src, void* dst, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VFLIP_H */
