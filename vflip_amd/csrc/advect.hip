// Semi-Lagrangian advection: RK2 midpoint backtrace with trilinear or clamped
// cubic gather (FP64, bit-exact with the reference arithmetic; trilinear also in
// FP32 for the simulation step), plus the smoke source/buoyancy forcing of the
// plume workload.
//
// Restates SubDomainSolver::advect_field (S/sub_domain.cpp:191-235, trilinear)
// and SmokeSolver3D::advect_field_cubic_clamp (S/Smoke_solver3D.cpp:657-707),
// with the Buffer3D samplers (S/fluid_buffer3D.h:205-356): interpolation in
// FP64, index clamping at the grid edges, positions rounded to float where the
// reference stores them in float. Compiled without FMA contraction so results
// are bit-identical with the reference arithmetic.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdlib>

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

// FP32 form of the trilinear RK2 advection (mode 2), for the simulation step:
// the same backtrace, clamped corners and interpolation order as k_advect_rk2
// mode 0, evaluated in FP32 with fused lerps a + c (b - a). Agrees with the FP64
// form to a few FP32 ulps. Each thread advects two cells of one (i, j) column
// (planes k, k+1) so the three dependent gather stages of the two cells overlap.
struct GridF {
    const float* f;
    int nx, ny, nz;
    float ox, oy, oz;
};

__device__ inline float lerpf(float a, float b, float c) { return __builtin_fmaf(c, b - a, a); }

// Grid coordinate w/h + o split into its cell index and the fraction inside the
// cell. q = w * (1/h) carries the FP32 rounding of the quotient; the exact
// remainder w - q h (one FMA) gives it back, so the fraction is accurate to
// ~1e-7 of a cell at any index (a plain FP32 w/h + o loses ~n * 6e-8 cells).
// POW2 (h a power of two): q is exact, the remainder is exactly 0 and g - floor(g)
// lies in [0, 1), so the general form reduces to one FMA, floor and difference
// with the same result bits.
template <bool POW2>
__device__ inline void cell_coord(float w, float h, float inv_h, float o, int& i, float& f) {
    if constexpr (POW2) {
        const float g = __builtin_fmaf(w, inv_h, o);
        const float fl = floorf(g);
        f = g - fl;
        i = (int)fl;
    } else {
        const float q = w * inv_h;
        const float c = __builtin_fmaf(-q, h, w) * inv_h;
        const float g = q + o;  // o is 0 or 0.5: exact for |q| < 2^22
        const float fl = floorf(g);
        f = (g - fl) + c;
        i = (int)fl;
        if (f < 0.f) { f += 1.f; --i; }
        else if (f >= 1.f) { f -= 1.f; ++i; }
    }
}

// The two clamped lattice lines around a coordinate along one axis, as byte
// offsets (index * stride): the corner loads are one 32-bit add each from a
// uniform base. stride < 2^24 and n * stride < 2^32 (checked by the entry points).
template <bool POW2>
__device__ inline void axis_lines(float w, float h, float inv_h, float o, int n, unsigned stride, unsigned& a0,
                                  unsigned& a1, float& f) {
    int i;
    cell_coord<POW2>(w, h, inv_h, o, i, f);
    a0 = __umul24((unsigned)min(max(i, 0), n - 1), stride);
    a1 = __umul24((unsigned)min(max(i + 1, 0), n - 1), stride);
}

__device__ inline float ldb(const float* base, unsigned off) {
    return *(const float*)((const char*)base + off);
}

// Trilinear sample in the order of the FP64 sampler: x lerps, then y, then z,
// with every corner index clamped to the grid.
template <bool POW2>
__device__ inline float sample_linear_f(const GridF& g, float h, float inv_h, float wx, float wy, float wz) {
    unsigned x0, x1, y0, y1, z0, z1;
    float cx, cy, cz;
    axis_lines<POW2>(wx, h, inv_h, g.ox, g.nx, 4u, x0, x1, cx);
    axis_lines<POW2>(wy, h, inv_h, g.oy, g.ny, 4u * g.nx, y0, y1, cy);
    axis_lines<POW2>(wz, h, inv_h, g.oz, g.nz, 4u * g.nx * g.ny, z0, z1, cz);
    const unsigned p00 = y0 + z0, p10 = y1 + z0, p01 = y0 + z1, p11 = y1 + z1;
    const float v1 = lerpf(lerpf(ldb(g.f, p00 + x0), ldb(g.f, p00 + x1), cx),
                           lerpf(ldb(g.f, p10 + x0), ldb(g.f, p10 + x1), cx), cy);
    const float v2 = lerpf(lerpf(ldb(g.f, p01 + x0), ldb(g.f, p01 + x1), cx),
                           lerpf(ldb(g.f, p11 + x0), ldb(g.f, p11 + x1), cx), cy);
    return lerpf(v1, v2, cz);
}

// The same sample where the axes flagged AX/AY/AZ are known to have a zero
// fraction (the position is on the grid's own lattice along them): the upper
// neighbour along such an axis is not loaded. lerpf(a, b, 0) == a for finite b
// (up to the sign of a zero), so the value equals sample_linear_f's.
template <bool AX, bool AY, bool AZ>
__device__ inline float sample_lattice_f(const GridF& g, float h, float inv_h, float wx, float wy, float wz) {
    unsigned x0, x1, y0, y1, z0, z1;
    float cx, cy, cz;
    axis_lines<true>(wx, h, inv_h, g.ox, g.nx, 4u, x0, x1, cx);
    axis_lines<true>(wy, h, inv_h, g.oy, g.ny, 4u * g.nx, y0, y1, cy);
    axis_lines<true>(wz, h, inv_h, g.oz, g.nz, 4u * g.nx * g.ny, z0, z1, cz);
    auto row = [&](unsigned p) {
        const float a = ldb(g.f, p + x0);
        if constexpr (AX) return a;
        else return lerpf(a, ldb(g.f, p + x1), cx);
    };
    auto plane = [&](unsigned z) {
        const float a = row(y0 + z);
        if constexpr (AY) return a;
        else return lerpf(a, row(y1 + z), cy);
    };
    const float v1 = plane(z0);
    if constexpr (AZ) return v1;
    else return lerpf(v1, plane(z1), cz);
}

// First RK2 stage at the output sample's own position: with a power-of-two h
// the position is exact, so along the axes where a velocity component's grid
// offset equals the field's the fraction is exactly 0. FIELD: 0 u faces,
// 1 v faces, 2 w faces, 3 cell centres; -1 general (all eight corners; POW2 as
// in cell_coord).
template <int FIELD, bool POW2>
__device__ inline void first_stage(const GridF& gu, const GridF& gv, const GridF& gw, float h, float inv_h, float x,
                                   float y, float z, float& u, float& v, float& w) {
    if constexpr (FIELD < 0) {
        u = sample_linear_f<POW2>(gu, h, inv_h, x, y, z);
        v = sample_linear_f<POW2>(gv, h, inv_h, x, y, z);
        w = sample_linear_f<POW2>(gw, h, inv_h, x, y, z);
    } else {
        // field offsets (FIELD == axis: 0.5 on that axis); component c has 0.5 on axis c
        constexpr bool fx = FIELD == 0, fy = FIELD == 1, fz = FIELD == 2;
        u = sample_lattice_f<fx == true, fy == false, fz == false>(gu, h, inv_h, x, y, z);
        v = sample_lattice_f<fx == false, fy == true, fz == false>(gv, h, inv_h, x, y, z);
        w = sample_lattice_f<fx == false, fy == false, fz == true>(gw, h, inv_h, x, y, z);
    }
}

template <int FIELD, bool POW2>
__global__ __launch_bounds__(kBlock) void k_advect_rk2_f32(int ni, int nj, int nk, float h, float inv_h, float dt,
                                                           const float* __restrict__ u, const float* __restrict__ v,
                                                           const float* __restrict__ w, GridF f0,
                                                           float* __restrict__ out0, GridF f1,
                                                           float* __restrict__ out1) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, kb = blockIdx.z * 2;
    if (i >= f0.nx || j >= f0.ny) return;
    const GridF gu{u, ni + 1, nj, nk, 0.5f, 0.f, 0.f};
    const GridF gv{v, ni, nj + 1, nk, 0.f, 0.5f, 0.f};
    const GridF gw{w, ni, nj, nk + 1, 0.f, 0.f, 0.5f};
    const float hdt = 0.5f * dt;
    float wx[2], wy[2], wz[2], uu[2], vv[2], ww[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int k = min(kb + c, f0.nz - 1);
        wx[c] = ((float)i - f0.ox) * h;
        wy[c] = ((float)j - f0.oy) * h;
        wz[c] = ((float)k - f0.oz) * h;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if constexpr (FIELD >= 0) first_stage_idx<FIELD>(gu, gv, gw, i, j, min(kb + c, f0.nz - 1), uu[c], vv[c], ww[c]);
        else first_stage<FIELD, POW2>(gu, gv, gw, h, inv_h, wx[c], wy[c], wz[c], uu[c], vv[c], ww[c]);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const float px = __builtin_fmaf(-hdt, uu[c], wx[c]), py = __builtin_fmaf(-hdt, vv[c], wy[c]);
        const float pz = __builtin_fmaf(-hdt, ww[c], wz[c]);
        uu[c] = sample_linear_f<POW2>(gu, h, inv_h, px, py, pz);
        vv[c] = sample_linear_f<POW2>(gv, h, inv_h, px, py, pz);
        ww[c] = sample_linear_f<POW2>(gw, h, inv_h, px, py, pz);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int k = kb + c;
        if (k >= f0.nz) break;
        // the reference's float ops: product, then difference
        const float px = wx[c] - dt * uu[c], py = wy[c] - dt * vv[c], pz = wz[c] - dt * ww[c];
        const int o = (k * f0.ny + j) * f0.nx + i;
        out0[o] = sample_linear_f<POW2>(f0, h, inv_h, px, py, pz);
        if (out1) out1[o] = sample_linear_f<POW2>(f1, h, inv_h, px, py, pz);
    }
}

// ----------------------------------------------------------------------------
// The plume step's five advections in one launch (mode 2 arithmetic): each thread
// a pair of cells of one column (planes k, k+1), producing their u, v, w faces and
// their rho, T centres (the faces i = ni, j = nj, k = nk come from the last cell
// of each row). Every value is the one k_advect_rk2_f32 computes (same samplers).
// ----------------------------------------------------------------------------
// first_stage<FIELD, true> with the lattice point's integer indices: the output
// sample of field FIELD at index (i, j, k) of its own grid (u face i, v face j, w
// face k, or cell centre) sits on each velocity component's lattice or halfway
// between two of its lines, so every fraction is 0 or 0.5 and every clamp is known;
// the loads and lerps are sample_lattice_f's in its order (same values, no
// coordinate arithmetic). ni = gv.nx, nj = gu.ny, nk = gu.nz.
template <int FIELD>
__device__ inline void first_stage_idx(const GridF& gu, const GridF& gv, const GridF& gw, int i, int j, int k,
                                       float& u, float& v, float& w) {
    const int ni = gv.nx, nj = gu.ny, nk = gu.nz;
    // 32-bit byte offsets from the uniform base, as the samplers (indices are in range)
    auto U = [&](int a, int b, int c) { return ldb(gu.f, 4u * (unsigned)((c * gu.ny + b) * gu.nx + a)); };
    auto Vv = [&](int a, int b, int c) { return ldb(gv.f, 4u * (unsigned)((c * gv.ny + b) * gv.nx + a)); };
    auto W = [&](int a, int b, int c) { return ldb(gw.f, 4u * (unsigned)((c * gw.ny + b) * gw.nx + a)); };
    if constexpr (FIELD == 3) {  // cell centre (i, j, k)
        u = lerpf(U(i, j, k), U(i + 1, j, k), 0.5f);
        v = lerpf(Vv(i, j, k), Vv(i, j + 1, k), 0.5f);
        w = lerpf(W(i, j, k), W(i, j, k + 1), 0.5f);
    } else if constexpr (FIELD == 0) {  // u face i of row j, plane k
        const int i0 = max(i - 1, 0), i1 = min(i, ni - 1);
        u = U(i, j, k);
        v = lerpf(lerpf(Vv(i0, j, k), Vv(i1, j, k), 0.5f), lerpf(Vv(i0, j + 1, k), Vv(i1, j + 1, k), 0.5f), 0.5f);
        w = lerpf(lerpf(W(i0, j, k), W(i1, j, k), 0.5f), lerpf(W(i0, j, k + 1), W(i1, j, k + 1), 0.5f), 0.5f);
    } else if constexpr (FIELD == 1) {  // v face j of column i, plane k
        const int j0 = max(j - 1, 0), j1 = min(j, nj - 1);
        u = lerpf(lerpf(U(i, j0, k), U(i + 1, j0, k), 0.5f), lerpf(U(i, j1, k), U(i + 1, j1, k), 0.5f), 0.5f);
        v = Vv(i, j, k);
        w = lerpf(lerpf(W(i, j0, k), W(i, j1, k), 0.5f), lerpf(W(i, j0, k + 1), W(i, j1, k + 1), 0.5f), 0.5f);
    } else {  // w face k of column i, row j
        const int k0 = max(k - 1, 0), k1 = min(k, nk - 1);
        u = lerpf(lerpf(U(i, j, k0), U(i + 1, j, k0), 0.5f), lerpf(U(i, j, k1), U(i + 1, j, k1), 0.5f), 0.5f);
        v = lerpf(lerpf(Vv(i, j, k0), Vv(i, j + 1, k0), 0.5f), lerpf(Vv(i, j, k1), Vv(i, j + 1, k1), 0.5f), 0.5f);
        w = W(i, j, k);
    }
}

// RK2 backtrace of the sample at (x, y, z): k_advect_rk2_f32's stages (FIELD as
// first_stage; LATTICE: its lattice form, valid for a power-of-two h).
template <int FIELD, bool LATTICE>
__device__ inline void backtrace(const GridF& gu, const GridF& gv, const GridF& gw, float h, float inv_h, float dt,
                                 float x, float y, float z, int i, int j, int k, float& px, float& py, float& pz) {
    float u, v, w;
    if constexpr (LATTICE) first_stage_idx<FIELD>(gu, gv, gw, i, j, k, u, v, w);
    else first_stage<-1, false>(gu, gv, gw, h, inv_h, x, y, z, u, v, w);
    const float hdt = 0.5f * dt;
    const float mx = __builtin_fmaf(-hdt, u, x), my = __builtin_fmaf(-hdt, v, y), mz = __builtin_fmaf(-hdt, w, z);
    u = sample_linear_f<LATTICE>(gu, h, inv_h, mx, my, mz);
    v = sample_linear_f<LATTICE>(gv, h, inv_h, mx, my, mz);
    w = sample_linear_f<LATTICE>(gw, h, inv_h, mx, my, mz);
    px = x - dt * u;
    py = y - dt * v;
    pz = z - dt * w;
}

template <bool LATTICE>
__global__ __launch_bounds__(256) void k_advect_plume(int ni, int nj, int nk, float h, float inv_h, float dt,
                                                      const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, const float* __restrict__ rho,
                                                      const float* __restrict__ temp, float* __restrict__ u2,
                                                      float* __restrict__ v2, float* __restrict__ w2,
                                                      float* __restrict__ rho2, float* __restrict__ temp2) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, kb = blockIdx.z * 2;
    if (i >= ni || j >= nj) return;
    const GridF gu{u, ni + 1, nj, nk, 0.5f, 0.f, 0.f}, gv{v, ni, nj + 1, nk, 0.f, 0.5f, 0.f},
        gw{w, ni, nj, nk + 1, 0.f, 0.f, 0.5f};
    const GridF gr{rho, ni, nj, nk, 0.f, 0.f, 0.f}, gt{temp, ni, nj, nk, 0.f, 0.f, 0.f};
    for (int dk = 0; dk < 2; ++dk) {
        const int k = kb + dk;
        if (k >= nk) break;
        const float x = (float)i * h, y = (float)j * h, z = (float)k * h;
        float px, py, pz;
        backtrace<3, LATTICE>(gu, gv, gw, h, inv_h, dt, x, y, z, i, j, k, px, py, pz);
        const int c = (k * nj + j) * ni + i;
        rho2[c] = sample_linear_f<LATTICE>(gr, h, inv_h, px, py, pz);
        temp2[c] = sample_linear_f<LATTICE>(gt, h, inv_h, px, py, pz);
        for (int q = 0; q < 1 + (i == ni - 1); ++q) {
            const int fi = i + q;
            backtrace<0, LATTICE>(gu, gv, gw, h, inv_h, dt, ((float)fi - 0.5f) * h, y, z, fi, j, k, px, py, pz);
            u2[(k * nj + j) * (ni + 1) + fi] = sample_linear_f<LATTICE>(gu, h, inv_h, px, py, pz);
        }
        for (int q = 0; q < 1 + (j == nj - 1); ++q) {
            const int fj = j + q;
            backtrace<1, LATTICE>(gu, gv, gw, h, inv_h, dt, x, ((float)fj - 0.5f) * h, z, i, fj, k, px, py, pz);
            v2[(k * (nj + 1) + fj) * ni + i] = sample_linear_f<LATTICE>(gv, h, inv_h, px, py, pz);
        }
        for (int q = 0; q < 1 + (k == nk - 1); ++q) {
            const int fk = k + q;
            backtrace<2, LATTICE>(gu, gv, gw, h, inv_h, dt, x, y, ((float)fk - 0.5f) * h, i, j, fk, px, py, pz);
            w2[(fk * nj + j) * ni + i] = sample_linear_f<LATTICE>(gw, h, inv_h, px, py, pz);
        }
    }
}

// Plume forcing on the reference-convention grid (cell centre (i,j,k)h):
//   source sphere: rho = T = 1 + jitter(seed, cell) inside |x - c| <= r
//   buoyancy on v faces: v += dt * (-alpha rho_f + beta T_f), rho_f = average of
//   the two cells sharing the face (faces 1..nj-1).
__global__ __launch_bounds__(kBlock) void k_plume_source(int ni, int nj, int nk, float h, float cx, float cy, float cz,
                                                         float r, unsigned seed, float jitter, float* __restrict__ rho,
                                                         float* __restrict__ temp) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= ni || j >= nj || k >= nk) return;
    const float dx = i * h - cx, dy = j * h - cy, dz = k * h - cz;
    if (dx * dx + dy * dy + dz * dz > r * r) return;
    const size_t c = ((size_t)k * nj + j) * ni + i;
    const float u = (float)randhash(seed * 0x9E3779B1u + (unsigned)c) * (1.0f / 4294967295.0f);
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

namespace {
// Every MAC component of an ni x nj x nk grid within the FP32 samplers' byte
// offsets: fewer than 2^30 samples (offsets < 2^32) and 2^22 per plane (the
// plane stride in bytes < 2^24 for the 24-bit multiply).
bool f32_addressable(int ni, int nj, int nk) {
    return (size_t)(ni + 1) * (nj + 1) * (nk + 1) < (1u << 30) && (size_t)(ni + 1) * (nj + 1) < (1u << 22);
}
}  // namespace

extern "C" {

int vflip_advect_rk2(int ni, int nj, int nk, float h, float dt, int mode, const float* u, const float* v,
                     const float* w, const float* field, int fnx, int fny, int fnz, float fox, float foy, float foz,
                     float* out, const float* field2, float* out2, void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && fnx > 0 && fny > 0 && fnz > 0, "grid dims");
    VF_ARG(u && v && w && field && out, "advection arrays");
    VF_ARG(mode == 0 || mode == 1 || mode == 2, "mode");
    VF_ARG((field2 == nullptr) == (out2 == nullptr), "second field and its output go together");
    VF_ARG((size_t)fnx * fny * fnz < (1u << 31) && (size_t)(ni + 1) * (nj + 1) * (nk + 1) < (1u << 31), "grid size");
    if (mode == 2) {
        // the FP32 samplers address corners by 32-bit byte offsets with 24-bit strides
        VF_ARG(f32_addressable(ni, nj, nk) && (size_t)fnx * fny * fnz < (1u << 30) && (size_t)fnx * fny < (1u << 22),
               "grid too large for mode 2 (at most 2^30 samples per grid, 2^22 per plane)");
        GridF g0{field, fnx, fny, fnz, fox, foy, foz};
        GridF g1{field2 ? field2 : field, fnx, fny, fnz, fox, foy, foz};
        dim3 grid((fnx + 63) / 64, (fny + 3) / 4, (fnz + 1) / 2);
        // lattice form of the first stage: h a power of two (positions and 1/h exact)
        // and the field on one of the MAC grid's four lattices
        int e2 = 0;
        const bool pow2 = std::frexp(h, &e2) == 0.5f;
        int field_kind = -1;
        if (pow2) {
            if (fox == 0.5f && foy == 0.f && foz == 0.f) field_kind = 0;
            else if (fox == 0.f && foy == 0.5f && foz == 0.f) field_kind = 1;
            else if (fox == 0.f && foy == 0.f && foz == 0.5f) field_kind = 2;
            else if (fox == 0.f && foy == 0.f && foz == 0.f) field_kind = 3;
        }
        const float inv_h = 1.0f / h;
        hipStream_t st = (hipStream_t)stream;
        switch (field_kind) {
            case 0: k_advect_rk2_f32<0, true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2); break;
            case 1: k_advect_rk2_f32<1, true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2); break;
            case 2: k_advect_rk2_f32<2, true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2); break;
            case 3: k_advect_rk2_f32<3, true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2); break;
            default:
                if (pow2)
                    k_advect_rk2_f32<-1, true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2);
                else
                    k_advect_rk2_f32<-1, false><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, inv_h, dt, u, v, w, g0, out, g1, out2);
        }
        VF_CHECK_LAUNCH();
        return 0;
    }
    Grid g0{field, fnx, fny, fnz, fox, foy, foz};
    Grid g1{field2 ? field2 : field, fnx, fny, fnz, fox, foy, foz};
    dim3 grid((fnx + 63) / 64, (fny + 3) / 4, fnz);
    k_advect_rk2<<<grid, dim3(64, 4), 0, (hipStream_t)stream>>>(ni, nj, nk, h, dt, mode, u, v, w, g0, out, g1, out2);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_advect_plume(int ni, int nj, int nk, float h, float dt, const float* u, const float* v, const float* w,
                       const float* rho, const float* temp, float* u2, float* v2, float* w2, float* rho2, float* temp2,
                       void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0, "grid dims");
    VF_ARG(u && v && w && rho && temp && u2 && v2 && w2 && rho2 && temp2, "advection arrays");
    VF_ARG(u2 != u && v2 != v && w2 != w && rho2 != rho && temp2 != temp, "outputs must not alias the inputs");
    VF_ARG(f32_addressable(ni, nj, nk), "grid too large (at most 2^30 samples per MAC grid, 2^22 per plane)");
    int e2 = 0;
    const bool pow2 = std::frexp(h, &e2) == 0.5f;
    const dim3 grid((ni + 63) / 64, (nj + 3) / 4, (nk + 1) / 2);
    hipStream_t st = (hipStream_t)stream;
    if (pow2)
        k_advect_plume<true><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, 1.0f / h, dt, u, v, w, rho, temp, u2, v2, w2,
                                                           rho2, temp2);
    else
        k_advect_plume<false><<<grid, dim3(64, 4), 0, st>>>(ni, nj, nk, h, 1.0f / h, dt, u, v, w, rho, temp, u2, v2,
                                                            w2, rho2, temp2);
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
