// The coarse solver's samplers with one subdomain (SmokeSolver3D::get_velocity,
// get_solid_phi, get_solid_normal, get_solid_vel; S/Smoke_solver3D.cpp:60-124,
// S/sub_domain.cpp:60-105), shared by the nested FLIP coupling and the tracers.
// A subdomain view with bmin = +inf and bmax = -inf (no_subdomain()) is never
// entered, so the same code samples the coarse grids alone. The including
// translation unit compiles without FMA contraction.
#pragma once

#include <cmath>

#include "common.h"
#include "sampling.h"

namespace vflip {

inline vflip_subdomain_view no_subdomain() {
    vflip_subdomain_view S{};
    S.ni = S.nj = S.nk = 1;
    S.h = 1.0f;
    for (int a = 0; a < 3; ++a) {
        S.bmin[a] = INFINITY;
        S.bmax[a] = -INFINITY;
    }
    return S;
}

// MAC samples of (u, v, w) on an nx x ny x nz grid of spacing h; NULL = 0.
__device__ inline void mac_sample(const float* u, const float* v, const float* w, int nx, int ny, int nz, float h,
                                  const float* x, float* o) {
    const double hd = h;
    o[0] = u ? sample_linear(Grid{u, nx + 1, ny, nz, 0.5, 0.0, 0.0}, hd, x[0], x[1], x[2]) : 0.f;
    o[1] = v ? sample_linear(Grid{v, nx, ny + 1, nz, 0.0, 0.5, 0.0}, hd, x[0], x[1], x[2]) : 0.f;
    o[2] = w ? sample_linear(Grid{w, nx, ny, nz + 1, 0.0, 0.0, 0.5}, hd, x[0], x[1], x[2]) : 0.f;
}

__device__ inline void sub_local(const vflip_subdomain_view& S, const float* x, float* q) {
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = x[a] - S.bmin[a];
}

// is_in (m = 2) / is_in2 (m = 5): bmin + m h < pos < bmax - m h, in double.
__device__ inline bool sub_inside(const vflip_subdomain_view& S, const float* x, double m) {
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) in = in && x[a] > S.bmin[a] + m * S.h && x[a] < S.bmax[a] - m * S.h;
    return in;
}

// get_boundary_coef: clamp(d / 4h, 0, 1), d = the least distance from pos to the
// faces of the box shrunk by 5h (float).
__device__ inline float boundary_coef(const vflip_subdomain_view& S, const float* x) {
    const float e = 5.0f * S.h;
    float d = 0.f;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        float m = 0.f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float b = side == 0 ? S.bmin[a] + e : S.bmax[a] - e;
            const float t = fabsf(x[a] - b);
            if (a == 0 || t < m) m = t;
        }
        d = side == 0 ? m : (m < d ? m : d);
    }
    const float c = d / (4.0f * S.h);
    return c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
}

// SmokeSolver3D::get_velocity with the subdomain blend.
__device__ inline void nested_velocity(const vflip_coarse_view& C, const vflip_subdomain_view& S, const float* x,
                                       float* o) {
    mac_sample(C.u, C.v, C.w, C.ni, C.nj, C.nk, C.h, x, o);
    if (sub_inside(S, x, 5.0)) {
        float q[3], o2[3];
        sub_local(S, x, q);
        mac_sample(S.u, S.v, S.w, S.ni, S.nj, S.nk, S.h, q, o2);
        const float c = boundary_coef(S, x);
#pragma unroll
        for (int a = 0; a < 3; ++a) o[a] = c * o2[a] + (1.0f - c) * o[a];
    }
}

__device__ inline float sub_phi(const vflip_subdomain_view& S, const float* q) {
    return sample_linear(Grid{S.solid_phi, S.ni + 1, S.nj + 1, S.nk + 1, 0.5, 0.5, 0.5}, (double)S.h, q[0], q[1],
                         q[2]);
}

// get_solid_phi: the subdomain's phi inside is_in, else the coarse grid's.
__device__ inline float nested_phi(const vflip_coarse_view& C, const vflip_subdomain_view& S, const float* x) {
    if (sub_inside(S, x, 2.0)) {
        float q[3];
        sub_local(S, x, q);
        return sub_phi(S, q);
    }
    return sample_linear(Grid{C.solid_phi, C.ni + 1, C.nj + 1, C.nk + 1, 0.5, 0.5, 0.5}, (double)C.h, x[0], x[1],
                         x[2]);
}

// get_solid_normal: normalize(sample_grad) of the same phi (S/vec.h:194-203,
// 218-219: a / (float)(sqrt((double)|a|^2) + 1e-6)).
__device__ inline void nested_normal(const vflip_coarse_view& C, const vflip_subdomain_view& S, const float* x,
                                     float* g) {
    if (sub_inside(S, x, 2.0)) {
        float q[3];
        sub_local(S, x, q);
        sample_grad(Grid{S.solid_phi, S.ni + 1, S.nj + 1, S.nk + 1, 0.5, 0.5, 0.5}, (double)S.h, q[0], q[1], q[2], g);
    } else {
        sample_grad(Grid{C.solid_phi, C.ni + 1, C.nj + 1, C.nk + 1, 0.5, 0.5, 0.5}, (double)C.h, x[0], x[1], x[2], g);
    }
    float l = g[0] * g[0];
    l += g[1] * g[1];
    l += g[2] * g[2];
    const float m = (float)(sqrt((double)l) + 1e-6);
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] /= m;
}

// get_solid_vel: the subdomain's solid velocity inside is_in, else the coarse grid's.
__device__ inline void nested_solid_vel(const vflip_coarse_view& C, const vflip_subdomain_view& S, const float* x,
                                        float* o) {
    if (sub_inside(S, x, 2.0)) {
        float q[3];
        sub_local(S, x, q);
        mac_sample(S.us, S.vs, S.ws, S.ni, S.nj, S.nk, S.h, q, o);
    } else {
        mac_sample(C.us, C.vs, C.ws, C.ni, C.nj, C.nk, C.h, x, o);
    }
}

}  // namespace vflip
