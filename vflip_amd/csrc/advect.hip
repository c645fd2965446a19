// Semi-Lagrangian advection: RK2 midpoint backtrace with trilinear or clamped
// cubic gather, plus the smoke source/buoyancy forcing of the plume workload.
//
// Restates SubDomainSolver::advect_field (S/sub_domain.cpp:191-235, trilinear)
// and SmokeSolver3D::advect_field_cubic_clamp (S/Smoke_solver3D.cpp:657-707),
// with the Buffer3D samplers (S/fluid_buffer3D.h:205-356): interpolation in
// FP64, index clamping at the grid edges, positions rounded to float where the
// reference stores them in float. Compiled without FMA contraction so results
// are bit-identical with the reference arithmetic.
#pragma clang fp contract(off)

#include "common.h"
#include "sampling.h"

namespace vflip {

// One field (or two fields on the same grid) advected by the MAC velocity.
__global__ __launch_bounds__(kBlock) void k_advect_rk2(int ni, int nj, int nk, float h, float dt, int mode,
                                                       const float* __restrict__ u, const float* __restrict__ v,
                                                       const float* __restrict__ w, Grid f0, float* __restrict__ out0,
                                                       Grid f1, float* __restrict__ out1) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= f0.nx || j >= f0.ny || k >= f0.nz) return;
    const double hd = h;
    const Grid gu{u, ni + 1, nj, nk, 0.5, 0.0, 0.0};
    const Grid gv{v, ni, nj + 1, nk, 0.0, 0.5, 0.0};
    const Grid gw{w, ni, nj, nk + 1, 0.0, 0.0, 0.5};
    const float wx = (float)(((double)(float)i - f0.ox) * hd);
    const float wy = (float)(((double)(float)j - f0.oy) * hd);
    const float wz = (float)(((double)(float)k - f0.oz) * hd);
    float uu = sample_linear(gu, hd, wx, wy, wz);
    float vv = sample_linear(gv, hd, wx, wy, wz);
    float ww = sample_linear(gw, hd, wx, wy, wz);
    float px = (float)((double)wx - 0.5 * (double)dt * (double)uu);
    float py = (float)((double)wy - 0.5 * (double)dt * (double)vv);
    float pz = (float)((double)wz - 0.5 * (double)dt * (double)ww);
    uu = sample_linear(gu, hd, px, py, pz);
    vv = sample_linear(gv, hd, px, py, pz);
    ww = sample_linear(gw, hd, px, py, pz);
    px = wx - dt * uu;
    py = wy - dt * vv;
    pz = wz - dt * ww;
    const size_t o = ((size_t)k * f0.ny + j) * f0.nx + i;
    for (int q = 0; q < 2; ++q) {
        const Grid& g = q == 0 ? f0 : f1;
        float* out = q == 0 ? out0 : out1;
        if (!out) continue;
        float val;
        if (mode == 0) {
            val = sample_linear(g, hd, px, py, pz);
        } else {
            // sample_cube_lerp corners + clamp to the trilinear value on overshoot
            const double gx = (double)px / hd + g.ox, gy = (double)py / hd + g.oy, gz = (double)pz / hd + g.oz;
            const int a = (int)floor(gx), b = (int)floor(gy), c = (int)floor(gz);
            float mn = at(g, a, b, c), mx = mn;
#pragma unroll
            for (int s = 1; s < 8; ++s) {
                const float t = at(g, a + (s & 1), b + ((s >> 1) & 1), c + ((s >> 2) & 1));
                mn = fminf(mn, t);
                mx = fmaxf(mx, t);
            }
            const float lin = sample_linear(g, hd, px, py, pz);
            const float cub = sample_cubic(g, hd, px, py, pz);
            val = (cub <= mn || cub >= mx) ? lin : cub;
        }
        out[o] = val;
    }
}

// Plume forcing on the reference-convention grid (cell centre (i,j,k)h):
//   source sphere: rho = T = 1 + jitter(seed, cell) inside |x - c| <= r
//   buoyancy on v faces: v += dt * (-alpha rho_f + beta T_f), rho_f = average of
//   the two cells sharing the face (faces 1..nj-1).
__device__ inline unsigned randhash_dev(unsigned seed) {
    unsigned i = (seed ^ 0xA3C59AC3u) * 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    return i;
}

__global__ __launch_bounds__(kBlock) void k_plume_source(int ni, int nj, int nk, float h, float cx, float cy, float cz,
                                                         float r, unsigned seed, float jitter, float* __restrict__ rho,
                                                         float* __restrict__ temp) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj || k >= nk) return;
    const float dx = i * h - cx, dy = j * h - cy, dz = k * h - cz;
    if (dx * dx + dy * dy + dz * dz > r * r) return;
    const size_t c = ((size_t)k * nj + j) * ni + i;
    const float u = (float)randhash_dev(seed * 0x9E3779B1u + (unsigned)c) * (1.0f / 4294967295.0f);
    rho[c] = 1.0f + jitter * u;
    temp[c] = 1.0f + jitter * u;
}

__global__ __launch_bounds__(kBlock) void k_buoyancy(int ni, int nj, int nk, float dt, float alpha, float beta,
                                                     const float* __restrict__ rho, const float* __restrict__ temp,
                                                     float* __restrict__ v) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j < 1 || j >= nj || k >= nk) return;
    const size_t c = ((size_t)k * nj + j) * ni + i, cm = c - ni;
    const float rf = 0.5f * (rho[c] + rho[cm]);
    const float tf = 0.5f * (temp[c] + temp[cm]);
    v[((size_t)k * (nj + 1) + j) * ni + i] += dt * (-alpha * rf + beta * tf);
}

}  // namespace vflip

using namespace vflip;

extern "C" {

int vflip_advect_rk2(int ni, int nj, int nk, float h, float dt, int mode, const float* u, const float* v,
                     const float* w, const float* field, int fnx, int fny, int fnz, float fox, float foy, float foz,
                     float* out, const float* field2, float* out2, void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && fnx > 0 && fny > 0 && fnz > 0, "grid dims");
    VF_ARG(u && v && w && field && out, "advection arrays");
    VF_ARG(mode == 0 || mode == 1, "mode");
    VF_ARG((field2 == nullptr) == (out2 == nullptr), "second field and its output go together");
    Grid g0{field, fnx, fny, fnz, fox, foy, foz};
    Grid g1{field2 ? field2 : field, fnx, fny, fnz, fox, foy, foz};
    dim3 grid((fnx + 63) / 64, (fny + 3) / 4, fnz);
    k_advect_rk2<<<grid, dim3(64, 4), 0, (hipStream_t)stream>>>(ni, nj, nk, h, dt, mode, u, v, w, g0, out, g1, out2);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_plume_force(int ni, int nj, int nk, float h, float dt, float cx, float cy, float cz, float r,
                      unsigned seed, float jitter, float alpha, float beta, float* rho, float* temp, float* v,
                      void* stream) {
    VF_ARG(ni > 0 && nj > 1 && nk > 0 && rho && temp && v, "plume arrays");
    hipStream_t s = (hipStream_t)stream;
    dim3 g((ni + 63) / 64, (nj + 3) / 4, nk);
    k_plume_source<<<g, dim3(64, 4), 0, s>>>(ni, nj, nk, h, cx, cy, cz, r, seed, jitter, rho, temp);
    k_buoyancy<<<g, dim3(64, 4), 0, s>>>(ni, nj, nk, dt, alpha, beta, rho, temp, v);
    VF_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
