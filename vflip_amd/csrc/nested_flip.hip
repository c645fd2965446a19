// Nested FLIP coupling: the coarse solver's particles and one subdomain.
//
// Restates, for device-resident particles (xyz-interleaved float pos/vel):
//   get_velocity (blend)  S/Smoke_solver3D.cpp:71-88, S/sub_domain.cpp:70-77, 96-105
//   get_solid_phi/normal/vel with the subdomain   S/Smoke_solver3D.cpp:60-124
//   advect_FLIP (coarse, nested)   S/Smoke_solver3D.cpp:141-154, 1408-1430
//   halo copy             S/Smoke_solver3D.cpp:393-401, is_halo S/sub_domain.h:203-214
//   seed_ghost            S/Smoke_solver3D.cpp:1529-1560
//   diffuse (PSE)         S/Smoke_solver3D.cpp:1562-1621, PSE_kernel S/Smoke_solver3D.h:204-209
//
// Samples interpolate in FP64 as Buffer3D does (sampling.h); compiled without FMA
// contraction so every float expression rounds as the reference's does.
#pragma clang fp contract(off)

#include <cmath>

#include "common.h"
#include "flip_internal.h"
#include "nested_sampling.h"

namespace vflip {

__global__ __launch_bounds__(kBlock) void k_nested_velocity(int n, const float* __restrict__ pos,
                                                            float* __restrict__ out, vflip_coarse_view C,
                                                            vflip_subdomain_view S) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float x[3] = {pos[3 * (size_t)p], pos[3 * (size_t)p + 1], pos[3 * (size_t)p + 2]};
    float o[3];
    nested_velocity(C, S, x, o);
#pragma unroll
    for (int a = 0; a < 3; ++a) out[3 * (size_t)p + a] = o[a];
}

// traceRK3 (coarse c3 = (1 - c1 - c2) dt) through the blended velocity, then
// pos -= phi * normal where phi < 0.
__global__ __launch_bounds__(kBlock) void k_advect_nested(int np, float* __restrict__ pos, float dt,
                                                          vflip_coarse_view C, vflip_subdomain_view S) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
    float x[3] = {pos[3 * (size_t)p], pos[3 * (size_t)p + 1], pos[3 * (size_t)p + 2]};
    float v1[3], v2[3], v3[3], m[3];
    nested_velocity(C, S, x, v1);
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = x[a] + t1 * v1[a];
    nested_velocity(C, S, m, v2);
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = x[a] + t2 * v2[a];
    nested_velocity(C, S, m, v3);
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = x[a] + c1 * v1[a] + c2 * v2[a] + c3 * v3[a];
    const float ph = nested_phi(C, S, x);
    if (ph < 0) {
        float g[3];
        nested_normal(C, S, x, g);
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] -= ph * g[a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) pos[3 * (size_t)p + a] = x[a];
}

// ------------------------------------------------------------------ halo copy
__device__ inline bool in_halo(const vflip_subdomain_view& S, float dx, const float* __restrict__ pos, size_t p) {
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = pos[3 * p + a];
        in = in && x > S.bmin[a] - dx && x < S.bmax[a] + dx;
    }
    return in;
}

__global__ __launch_bounds__(kBlock) void k_halo_flag(int np, const float* __restrict__ pos, vflip_subdomain_view S,
                                                      float dx, int* __restrict__ flag) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < np) flag[p] = in_halo(S, dx, pos, p);
}

// offset = the exclusive scan of the flags: particle p is selected when
// offset[p + 1] > offset[p], and goes to out[offset[p]].
__global__ __launch_bounds__(kBlock) void k_halo_emit(int np, const float* __restrict__ pos,
                                                      const float* __restrict__ vel, const int* __restrict__ offset,
                                                      float* __restrict__ out_pos, float* __restrict__ out_vel) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const int o = offset[p];
    if (offset[p + 1] == o) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        out_pos[3 * (size_t)o + a] = pos[3 * (size_t)p + a];
        out_vel[3 * (size_t)o + a] = vel[3 * (size_t)p + a];
    }
}

// ---------------------------------------------------------------- seed_ghost
// Candidate r of subdomain cell c = (i, j, k): (i,j,k) h + 0.49 h r + bmin with
// r = frand(-1, 1) per axis from the counter hash, key 3 (8 c + r) + axis.
__device__ inline void ghost_pos(const vflip_subdomain_view& S, unsigned seed, size_t c, int i, int j, int k, int r,
                                 float* x) {
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float jit = reseed_jitter(seed, (unsigned)(3 * (8 * c + (size_t)r) + a));
        x[a] = ((float)ijk[a] * S.h + (jit * 0.49f) * S.h) + S.bmin[a];
    }
}

__device__ inline bool ghost_kept(const vflip_subdomain_view& S, const float* x, float& phi) {
    float q[3];
    sub_local(S, x, q);
    phi = sub_phi(S, q);
    return phi <= 2.0 * S.h && phi > -2.0 * S.h;
}

__device__ inline void ghost_cell(const vflip_subdomain_view& S, size_t c, int& i, int& j, int& k) {
    const unsigned plane = (unsigned)S.ni * (unsigned)S.nj, cu = (unsigned)c;
    k = (int)(cu / plane);
    const int r = (int)(cu - (unsigned)k * plane);
    j = r / S.ni;
    i = r - j * S.ni;
}

// Per cell: the mask of kept candidates (cells whose get_solid_phi at
// bmin + (i,j,k) h is >= 3h make none) and its popcount.
__global__ __launch_bounds__(kBlock) void k_ghost_count(size_t ncell, vflip_coarse_view C, vflip_subdomain_view S,
                                                        unsigned seed, int* __restrict__ cnt, int* __restrict__ mask) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    int i, j, k;
    ghost_cell(S, c, i, j, k);
    const int ijk[3] = {i, j, k};
    float cp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) cp[a] = (float)ijk[a] * S.h + S.bmin[a];
    int m = 0;
    if (nested_phi(C, S, cp) < 3.0f * S.h) {
        for (int r = 0; r < 8; ++r) {
            float x[3], phi;
            ghost_pos(S, seed, c, i, j, k, r, x);
            if (ghost_kept(S, x, phi)) m |= 1 << r;
        }
    }
    cnt[c] = __popc(m);
    mask[c] = m;
}

__global__ __launch_bounds__(kBlock) void k_ghost_emit(size_t ncell, vflip_coarse_view C, vflip_subdomain_view S,
                                                       unsigned seed, const int* __restrict__ offset,
                                                       const int* __restrict__ mask, float* __restrict__ out_pos,
                                                       float* __restrict__ out_vel) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    const int m = mask[c];
    if (!m) return;
    int i, j, k;
    ghost_cell(S, c, i, j, k);
    size_t o = (size_t)offset[c];
    for (int r = 0; r < 8; ++r) {
        if (!((m >> r) & 1)) continue;
        float x[3], phi, vv[3];
        ghost_pos(S, seed, c, i, j, k, r, x);
        ghost_kept(S, x, phi);
        nested_velocity(C, S, x, vv);
        if (phi <= 0) nested_solid_vel(C, S, x, vv);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            out_pos[3 * o + a] = x[a];
            out_vel[3 * o + a] = vv[a];
        }
        ++o;
    }
}

// -------------------------------------------------------------------- diffuse
// PSE_kernel as written: (1/frac*frac*frac) exp(-|xi - xj|^2 0.25 / tau), the
// distance in float, the exponential in double.
__device__ inline float pse_kernel(const float* xi, const float* xj, float tau, float fr) {
    float d = (xi[0] - xj[0]) * (xi[0] - xj[0]);
    d += (xi[1] - xj[1]) * (xi[1] - xj[1]);
    d += (xi[2] - xj[2]) * (xi[2] - xj[2]);
    return (float)((double)fr * exp(-d * 0.25 / tau));
}

// du = sum (v_j - v_i) K / (1e-9 + sum K) over the particles binned in the
// 3x3x3 cells around particle i (z, y, x ascending, list order). The threads
// walk the bins: thread (cell c, slot s) takes the particles c's list holds at
// s, s + slots, ... and diffuses the real ones among them (clamped binning
// lists every particle once). A wave then covers 64 / slots neighbouring cells
// whose neighbour lists overlap, so the gathers of neighbour positions and
// velocities are shared through the caches instead of being fetched once per
// particle in arbitrary order. That pays where the bins are dense (at least
// kPseDense real particles per cell); sparse bins, like a halo copy of a
// coarse fill already in cell order, keep one thread per real particle
// (k_pse_diffuse_particles), which leaves no lane idle.
constexpr int kPseSlots = 8;
constexpr int kPseDense = 4;
__device__ inline void pse_one(int i, const float* __restrict__ pos, const float* __restrict__ vel,
                               const int* __restrict__ start, const int* __restrict__ order,
                               const vflip_coarse_view& C, const vflip_subdomain_view& S, float tau, float fr,
                               float* __restrict__ out_vel) {
    const float xi[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
    const float ui[3] = {vel[3 * (size_t)i], vel[3 * (size_t)i + 1], vel[3 * (size_t)i + 2]};
    int ijk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ijk[a] = trunc_i32((xi[a] - S.bmin[a]) / S.h + 0.5f);
    float du[3] = {0.f, 0.f, 0.f};
    if (nested_phi(C, S, xi) <= 2.0 * C.h && ijk[0] >= 0 && ijk[0] < S.ni && ijk[1] >= 0 && ijk[1] < S.nj &&
        ijk[2] >= 0 && ijk[2] < S.nk) {
        float coef = 1e-9f;
        float ue[3] = {0.f, 0.f, 0.f};
        for (int kk = max(ijk[2] - 1, 0); kk <= min(ijk[2] + 1, S.nk - 1); ++kk)
            for (int jj = max(ijk[1] - 1, 0); jj <= min(ijk[1] + 1, S.nj - 1); ++jj)
                for (int ii = max(ijk[0] - 1, 0); ii <= min(ijk[0] + 1, S.ni - 1); ++ii) {
                    const size_t cidx = ((size_t)kk * S.nj + jj) * S.ni + ii;
                    const int e = start[cidx + 1];
                    for (int q = start[cidx]; q < e; ++q) {
                        const size_t pj = (size_t)order[q];
                        const float xj[3] = {pos[3 * pj], pos[3 * pj + 1], pos[3 * pj + 2]};
                        const float K = pse_kernel(xi, xj, tau, fr);
#pragma unroll
                        for (int a = 0; a < 3; ++a) ue[a] += (vel[3 * pj + a] - ui[a]) * K;
                        coef += K;
                    }
                }
#pragma unroll
        for (int a = 0; a < 3; ++a) du[a] = ue[a] / coef;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) out_vel[3 * (size_t)i + a] = ui[a] + du[a];
}

__global__ __launch_bounds__(kBlock) void k_pse_diffuse_bins(int np_real, int ncell, const float* __restrict__ pos,
                                                             const float* __restrict__ vel,
                                                             const int* __restrict__ start,
                                                             const int* __restrict__ order, vflip_coarse_view C,
                                                             vflip_subdomain_view S, float tau, float fr,
                                                             float* __restrict__ out_vel) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int c = t / kPseSlots, slot = t % kPseSlots;
    if (c >= ncell) return;
    const int e = start[c + 1];
    for (int q = start[c] + slot; q < e; q += kPseSlots) {
        const int i = order[q];
        if (i < np_real) pse_one(i, pos, vel, start, order, C, S, tau, fr, out_vel);
    }
}

__global__ __launch_bounds__(kBlock) void k_pse_diffuse_particles(int np_real, const float* __restrict__ pos,
                                                                  const float* __restrict__ vel,
                                                                  const int* __restrict__ start,
                                                                  const int* __restrict__ order, vflip_coarse_view C,
                                                                  vflip_subdomain_view S, float tau, float fr,
                                                                  float* __restrict__ out_vel) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < np_real) pse_one(i, pos, vel, start, order, C, S, tau, fr, out_vel);
}

static bool view_ok(const vflip_coarse_view* C) {
    return C && C->ni > 0 && C->nj > 0 && C->nk > 0 && C->h > 0 && C->u && C->v && C->w && C->solid_phi;
}
static bool view_ok(const vflip_subdomain_view* S) {
    return S && S->ni > 0 && S->nj > 0 && S->nk > 0 && S->h > 0 && S->u && S->v && S->w && S->solid_phi;
}


}  // namespace vflip

using namespace vflip;

int vflip_nested_velocity(const vflip_coarse_view* C, const vflip_subdomain_view* S, int n, const float* pos,
                          float* out, void* hip_stream) {
    VF_ARG(view_ok(C) && view_ok(S) && n >= 0, "nested velocity args");
    VF_ARG(n == 0 || (pos && out), "nested velocity arrays");
    if (n > 0) k_nested_velocity<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(n, pos, out, *C, *S);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_advect_particles_nested(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, int np,
                                  float* pos, void* hip_stream) {
    VF_ARG(view_ok(C) && view_ok(S) && np >= 0, "nested advection args");
    VF_ARG(np == 0 || pos, "nested advection particles");
    if (np > 0) k_advect_nested<<<grid_for(np), kBlock, 0, (hipStream_t)hip_stream>>>(np, pos, dt, *C, *S);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_halo_select(vflip_particles* P, const vflip_subdomain_view* S, float dx, int np, const float* pos,
                      const float* vel, int cap, float* out_pos, float* out_vel, int* nout) {
    VF_ARG(P && S && nout && np >= 0 && cap >= 0, "halo select args");
    VF_ARG(np == 0 || (pos && vel), "halo select particles");
    VF_ARG(cap == 0 || (out_pos && out_vel), "halo select outputs");
    hipStream_t s = P->stream;
    if (int rc = ensure_selection_scratch(P, (size_t)np)) return rc;
    if (np > 0) k_halo_flag<<<grid_for(np), kBlock, 0, s>>>(np, pos, *S, dx, P->pflag);
    scan_exclusive_i32(P->pflag, (size_t)np, P->ppartial, s);
    VF_CHECK_LAUNCH();
    int total = 0;
    VF_HIP(hipMemcpyAsync(&total, P->pflag + np, sizeof(int), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    *nout = total;
    if (total > cap) {
        set_error("halo select: %d particles do not fit the output capacity %d", total, cap);
        return VFLIP_ERR_CAPACITY;
    }
    if (total > 0) k_halo_emit<<<grid_for(np), kBlock, 0, s>>>(np, pos, vel, P->pflag, out_pos, out_vel);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_seed_ghost(vflip_particles* P, const vflip_coarse_view* C, const vflip_subdomain_view* S, unsigned seed,
                     int cap, float* out_pos, float* out_vel, int* nout) {
    VF_ARG(P && view_ok(C) && view_ok(S) && nout && cap >= 0, "seed ghost args");
    VF_ARG(P->nx == S->ni && P->ny == S->nj && P->nz == S->nk, "seed ghost: P must be the subdomain's workspace");
    VF_ARG(cap == 0 || (out_pos && out_vel), "seed ghost outputs");
    hipStream_t s = P->stream;
    const size_t nc = P->ncell;
    if (!P->reo && hipMalloc(&P->reo, (2 * nc + 1) * sizeof(int)) != hipSuccess) {
        P->reo = nullptr;
        set_error("ghost workspace of %zu cells: out of device memory", nc);
        return VFLIP_ERR_NOMEM;
    }
    int* cnt = P->reo;
    int* mask = P->reo + nc + 1;
    k_ghost_count<<<grid_for(nc), kBlock, 0, s>>>(nc, *C, *S, seed, cnt, mask);
    scan_exclusive_i32(cnt, nc, P->partial, s);
    VF_CHECK_LAUNCH();
    int total = 0;
    VF_HIP(hipMemcpyAsync(&total, cnt + nc, sizeof(int), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    *nout = total;
    if (total > cap) {
        set_error("seed ghost: %d particles do not fit the output capacity %d", total, cap);
        return VFLIP_ERR_CAPACITY;
    }
    if (total > 0) k_ghost_emit<<<grid_for(nc), kBlock, 0, s>>>(nc, *C, *S, seed, cnt, mask, out_pos, out_vel);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_pse_diffuse(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, float nu, int np_real,
                      const float* pos, const float* vel, const int* cell_start, const int* order, float* out_vel,
                      void* hip_stream) {
    VF_ARG(view_ok(C) && view_ok(S) && np_real >= 0 && cell_start, "pse diffuse args");
    VF_ARG(np_real == 0 || (pos && vel && order && out_vel), "pse diffuse arrays");
    // tau = dt nu + 1e-10 passed as float; frac = (float)sqrt(4 pi tau) (the
    // reference's 3.1415926), and the kernel's factor 1/frac*frac*frac in float
    const float tau = (float)(dt * nu + 1e-10);
    const float frac = (float)std::sqrt(4.0 * 3.1415926 * tau);
    const float fr = 1.0f / frac * frac * frac;
    const size_t ncell = (size_t)S->ni * S->nj * S->nk;
    VF_ARG(ncell * kPseSlots < (1u << 31), "subdomain cell count");
    hipStream_t st = (hipStream_t)hip_stream;
    if (np_real > 0 && (size_t)np_real >= kPseDense * ncell)
        k_pse_diffuse_bins<<<grid_for(ncell * kPseSlots), kBlock, 0, st>>>(np_real, (int)ncell, pos, vel, cell_start,
                                                                           order, *C, *S, tau, fr, out_vel);
    else if (np_real > 0)
        k_pse_diffuse_particles<<<grid_for(np_real), kBlock, 0, st>>>(np_real, pos, vel, cell_start, order, *C, *S,
                                                                       tau, fr, out_vel);
    VF_CHECK_LAUNCH();
    return 0;
}
