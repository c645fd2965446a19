// FP64 samplers of the reference Buffer3D on linear (x-fastest) device grids.
//
// Buffer3D::at (index clamping), sample_linear, cubic_interp and sample_cubic,
// S/fluid_buffer3D.h:22-29, 205-326: grid coordinate (x/h + offset) in double,
// floor, clamped corner reads, lerps in double, result rounded to float. The
// including translation unit must compile with FMA contraction off for these
// to round exactly as the reference does.
#pragma once
#pragma clang fp contract(off)

#include "common.h"

namespace vflip {

struct Grid {
    const float* f;
    int nx, ny, nz;
    double ox, oy, oz;  // grid offsets in cells (u: .5,0,0)
};

__device__ inline float at(const Grid& g, int i, int j, int k) {
    i = i < 0 ? 0 : (i > g.nx - 1 ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j > g.ny - 1 ? g.ny - 1 : j);
    k = k < 0 ? 0 : (k > g.nz - 1 ? g.nz - 1 : k);
    return g.f[((size_t)k * g.ny + j) * g.nx + i];
}

__device__ inline double lerpd(double a, double b, double c) { return (1.0 - c) * a + c * b; }

// A grid whose values are new - old, each difference rounded to float as the
// reference's getDelta stores it (S/Smoke_solver3D.cpp:1297-1318).
struct DeltaGrid {
    const float* fnew;
    const float* fold;
    int nx, ny, nz;
    double ox, oy, oz;
};

__device__ inline float at(const DeltaGrid& g, int i, int j, int k) {
    i = i < 0 ? 0 : (i > g.nx - 1 ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j > g.ny - 1 ? g.ny - 1 : j);
    k = k < 0 ? 0 : (k > g.nz - 1 ? g.nz - 1 : k);
    const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
    return g.fnew[c] - g.fold[c];
}

// Value at element offset c (in range).
__device__ inline float val(const Grid& g, size_t c) { return g.f[c]; }
__device__ inline float val(const DeltaGrid& g, size_t c) { return g.fnew[c] - g.fold[c]; }

// Grid coordinates w / h + o of a sample. For a power-of-two h (the usual grid
// spacings) w / h is an exact scaling, the same value as w times the exact
// reciprocal, which takes one multiply instead of a division.
__device__ inline void grid_coords(double h, double wx, double wy, double wz, double ox, double oy, double oz,
                                   double& gx, double& gy, double& gz) {
    const unsigned long long hb = (unsigned long long)__double_as_longlong(h);
    const unsigned e = (unsigned)(hb >> 52) & 0x7FFu;
    if ((hb & 0x000FFFFFFFFFFFFFull) == 0 && e >= 1 && e <= 2045) {
        const double ih = __longlong_as_double((long long)(((unsigned long long)(2046 - e) << 52) |
                                                           (hb & 0x8000000000000000ull)));
        gx = wx * ih + ox;
        gy = wy * ih + oy;
        gz = wz * ih + oz;
    } else {
        gx = wx / h + ox;
        gy = wy / h + oy;
        gz = wz / h + oz;
    }
}

// Buffer3D::sample_linear: floor of the grid coordinate, corners clamped to the
// grid (Buffer3D::at), lerps in double in the order x, y, z. Each axis is clamped
// once for its two lattice lines and the eight corners are offsets from the four
// (y, z) rows (same loads and values as at()).
template <typename G>
__device__ inline float sample_linear(const G& g, double h, double wx, double wy, double wz) {
    double gx, gy, gz;
    grid_coords(h, wx, wy, wz, g.ox, g.oy, g.oz, gx, gy, gz);
    const int i = (int)floor(gx), j = (int)floor(gy), k = (int)floor(gz);
    const double cx = gx - (double)i, cy = gy - (double)j, cz = gz - (double)k;
    const int i0 = min(max(i, 0), g.nx - 1), i1 = min(max(i + 1, 0), g.nx - 1);
    const int j0 = min(max(j, 0), g.ny - 1), j1 = min(max(j + 1, 0), g.ny - 1);
    const int k0 = min(max(k, 0), g.nz - 1), k1 = min(max(k + 1, 0), g.nz - 1);
    const size_t nx = (size_t)g.nx, ny = (size_t)g.ny;
    const size_t r00 = ((size_t)k0 * ny + j0) * nx, r10 = ((size_t)k0 * ny + j1) * nx;
    const size_t r01 = ((size_t)k1 * ny + j0) * nx, r11 = ((size_t)k1 * ny + j1) * nx;
    const float c000 = val(g, r00 + i0), c100 = val(g, r00 + i1), c010 = val(g, r10 + i0), c110 = val(g, r10 + i1);
    const float c001 = val(g, r01 + i0), c101 = val(g, r01 + i1), c011 = val(g, r11 + i0), c111 = val(g, r11 + i1);
    const double v1 = lerpd(lerpd(c000, c100, cx), lerpd(c010, c110, cx), cy);
    const double v2 = lerpd(lerpd(c001, c101, cx), lerpd(c011, c111, cx), cy);
    return (float)lerpd(v1, v2, cz);
}

// Buffer3D::sample_grad (S/fluid_buffer3D.h:249-258): central differences at
// +-h taken in float, divided by h in double, rounded to float, times 0.5f.
__device__ inline void sample_grad(const Grid& g, double h, double wx, double wy, double wz, float* out) {
    const float right = sample_linear(g, h, wx + h, wy, wz), left = sample_linear(g, h, wx - h, wy, wz);
    const float up = sample_linear(g, h, wx, wy + h, wz), bottom = sample_linear(g, h, wx, wy - h, wz);
    const float front = sample_linear(g, h, wx, wy, wz - h), back = sample_linear(g, h, wx, wy, wz + h);
    out[0] = 0.5f * (float)((double)(right - left) / h);
    out[1] = 0.5f * (float)((double)(up - bottom) / h);
    out[2] = 0.5f * (float)((double)(back - front) / h);
}

// Buffer3D::cubic_interp with float weights (element type) and double values.
__device__ inline double cubic_interp(double m1, double v0, double v1, double v2, double fd) {
    const float f = (float)fd;
    const float f2 = f * f, f3 = f2 * f;
    const float wm1 = -float(1. / 3) * f + float(1. / 2) * f2 - float(1. / 6) * f3;
    const float w0 = 1 - f2 + float(1. / 2) * (f3 - f);
    const float w1 = f + float(1. / 2) * (f2 - f3);
    const float w2 = float(1. / 6) * (f3 - f);
    return wm1 * m1 + w0 * v0 + w1 * v1 + w2 * v2;
}

// Buffer3D::sample_cubic
__device__ inline float sample_cubic(const Grid& g, double h, double wx, double wy, double wz) {
    const double gx = wx / h + g.ox, gy = wy / h + g.oy, gz = wz / h + g.oz;
    const int i = (int)floor(gx), j = (int)floor(gy), k = (int)floor(gz);
    const double x = gx - (double)i, y = gy - (double)j, z = gz - (double)k;
    double fz[4];
#pragma unroll
    for (int dk = -1; dk <= 2; ++dk) {
        double fy[4];
#pragma unroll
        for (int dj = -1; dj <= 2; ++dj)
            fy[dj + 1] = cubic_interp(at(g, i - 1, j + dj, k + dk), at(g, i, j + dj, k + dk), at(g, i + 1, j + dj, k + dk),
                                      at(g, i + 2, j + dj, k + dk), x);
        fz[dk + 1] = cubic_interp(fy[0], fy[1], fy[2], fy[3], y);
    }
    return (float)cubic_interp(fz[0], fz[1], fz[2], fz[3], z);
}

}  // namespace vflip
