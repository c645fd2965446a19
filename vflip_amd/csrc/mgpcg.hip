// Multigrid-preconditioned CG pressure solver for gfx950.
//
// Restates the reference aggregation-AMG PCG (S/AlgebraicMultigrid.h:22-301,
// S/GeometricLevelGen.{h,cpp}) matrix-free: every level is a symmetric 7-point
// operator stored as diag + couplings to the +x/+y/+z neighbour (16 B/cell in
// FP32), built in closed form on the GPU instead of sparse triple products.
//
// Storage per level l (linear, x fastest):
//   d, ax, ay, az, leak  FP32 (leak = diag - sum|couplings|, the Dirichlet part,
//                        only read when building the next level)
//   hd, hax, hay, haz    FP16 copies scaled by 1/hscale (mixed precision, l >= 2)
//   x, b                 FP32 V-cycle vectors (level 0 uses the PCG r and z)
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace vflip {

// ----------------------------------------------------------------------------
// Operator view used by the kernels. CT = float (FP32 storage) or __half.
// ----------------------------------------------------------------------------
template <typename CT> struct Op {
    const CT* d;
    const CT* ax;
    const CT* ay;
    const CT* az;
    float scale;
    int ni, nj, nk;
    __device__ inline float D(size_t c) const { return scale * (float)d[c]; }
    __device__ inline float AX(size_t c) const { return scale * (float)ax[c]; }
    __device__ inline float AY(size_t c) const { return scale * (float)ay[c]; }
    __device__ inline float AZ(size_t c) const { return scale * (float)az[c]; }
};

__device__ inline void unflat(size_t c, int ni, int nj, int& i, int& j, int& k) {
    size_t t = c / (size_t)ni;
    i = (int)(c - t * ni);
    k = (int)(t / (size_t)nj);
    j = (int)(t - (size_t)k * nj);
}

// Off-diagonal row sum sum_{j != c} A(c,j) x_j in the reference CSR column order
// (-z, -y, -x, +x, +y, +z), S/AlgebraicMultigrid.h:41-53.
template <typename CT>
__device__ inline float offdiag_sum(const Op<CT>& op, const float* __restrict__ x, size_t c, int i, int j, int k) {
    const size_t sl = (size_t)op.ni * op.nj;
    float s = 0.f;
    if (k > 0) s += op.AZ(c - sl) * x[c - sl];
    if (j > 0) s += op.AY(c - op.ni) * x[c - op.ni];
    if (i > 0) s += op.AX(c - 1) * x[c - 1];
    if (i + 1 < op.ni) s += op.AX(c) * x[c + 1];
    if (j + 1 < op.nj) s += op.AY(c) * x[c + op.ni];
    if (k + 1 < op.nk) s += op.AZ(c) * x[c + sl];
    return s;
}

// ----------------------------------------------------------------------------
// Smoother: one red-black half sweep, colour (i+j+k)&1 == colour, in place.
// x = (b - offdiag)/diag, diag == 0 -> 0 (S/AlgebraicMultigrid.h:54-62).
// ----------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_half(Op<CT> op, const float* __restrict__ b, float* __restrict__ x,
                                                      int colour) {
    const size_t n = (size_t)op.ni * op.nj * op.nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat(c, op.ni, op.nj, i, j, k);
    if (((i + j + k) & 1) != colour) return;
    float s = offdiag_sum(op, x, c, i, j, k);
    float d = op.D(c);
    x[c] = (d != 0.f) ? (b[c] - s) / d : 0.f;
}

// First half sweep from a zero guess: odd cells get b/diag, even cells 0.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_first(Op<CT> op, const float* __restrict__ b, float* __restrict__ x) {
    const size_t n = (size_t)op.ni * op.nj * op.nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat(c, op.ni, op.nj, i, j, k);
    float v = 0.f;
    if ((i + j + k) & 1) {
        float d = op.D(c);
        v = (d != 0.f) ? (b[c] - 0.f) / d : 0.f;
    }
    x[c] = v;
}

// Odd half sweep fused with the preceding prolongation x += P xc: odd cells read
// their even neighbours as x + xc(parent) (S/AlgebraicMultigrid.h:118-129, 161-165).
// Even cells are left uncorrected: the following even half sweep overwrites them
// without reading their old value.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_odd_prolong(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, const float* __restrict__ xc,
                                                             int ci, int cj) {
    const size_t n = (size_t)op.ni * op.nj * op.nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat(c, op.ni, op.nj, i, j, k);
    if (!((i + j + k) & 1)) return;
    const size_t sl = (size_t)op.ni * op.nj;
    auto X = [&](size_t q, int a, int bb, int cc) { return x[q] + xc[lin3(a >> 1, bb >> 1, cc >> 1, ci, cj)]; };
    float s = 0.f;
    if (k > 0) s += op.AZ(c - sl) * X(c - sl, i, j, k - 1);
    if (j > 0) s += op.AY(c - op.ni) * X(c - op.ni, i, j - 1, k);
    if (i > 0) s += op.AX(c - 1) * X(c - 1, i - 1, j, k);
    if (i + 1 < op.ni) s += op.AX(c) * X(c + 1, i + 1, j, k);
    if (j + 1 < op.nj) s += op.AY(c) * X(c + op.ni, i, j + 1, k);
    if (k + 1 < op.nk) s += op.AZ(c) * X(c + sl, i, j, k + 1);
    float d = op.D(c);
    x[c] = (d != 0.f) ? (b[c] - s) / d : 0.f;
}

// x += P xc (injection), used when there are no post-smoothing sweeps.
__global__ __launch_bounds__(kBlock) void k_prolong_add(int ni, int nj, int nk, float* __restrict__ x,
                                                        const float* __restrict__ xc, int ci, int cj) {
    const size_t n = (size_t)ni * nj * nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat(c, ni, nj, i, j, k);
    x[c] += xc[lin3(i >> 1, j >> 1, k >> 1, ci, cj)];
}

// b_c = R (b - A x), R = 1/8 per existing child (S/AlgebraicMultigrid.h:105-117,
// S/GeometricLevelGen.cpp:50-63). One thread per coarse cell.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_residual_restrict(Op<CT> op, const float* __restrict__ b,
                                                              const float* __restrict__ x, float* __restrict__ bc,
                                                              int ci, int cj, int ck) {
    const size_t nc = (size_t)ci * cj * ck;
    size_t C = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (C >= nc) return;
    int I, J, K;
    unflat(C, ci, cj, I, J, K);
    float acc = 0.f;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                if (i >= op.ni || j >= op.nj || k >= op.nk) continue;
                size_t c = lin3(i, j, k, op.ni, op.nj);
                float r = b[c] - (offdiag_sum(op, x, c, i, j, k) + op.D(c) * x[c]);
                acc += 0.125f * r;
            }
    bc[C] = acc;
}

// ----------------------------------------------------------------------------
// Coarsest level: all RB-GS sweeps in one workgroup, x resident in LDS,
// coefficients in registers (n <= 4096 unknowns, S/GeometricLevelGen.h:46).
// ----------------------------------------------------------------------------
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseMax = 4096;
constexpr int kCoarsePer = kCoarseMax / kCoarseThreads;

template <typename CT>
__global__ __launch_bounds__(kCoarseThreads) void k_coarsest(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, int sweeps) {
    __shared__ float xs[kCoarseMax + 1];
    const int n = op.ni * op.nj * op.nk;
    const int sl = op.ni * op.nj;
    float cd[kCoarsePer], cb[kCoarsePer], ca[kCoarsePer][6];
    int nb[kCoarsePer][6];
    int colour[kCoarsePer];
#pragma unroll
    for (int q = 0; q < kCoarsePer; ++q) {
        int c = threadIdx.x + q * kCoarseThreads;
        colour[q] = -1;
        if (c < n) {
            int i = c % op.ni, t = c / op.ni, j = t % op.nj, k = t / op.nj;
            colour[q] = (i + j + k) & 1;
            cd[q] = op.D(c);
            cb[q] = b[c];
            nb[q][0] = k > 0 ? c - sl : kCoarseMax;
            ca[q][0] = k > 0 ? op.AZ(c - sl) : 0.f;
            nb[q][1] = j > 0 ? c - op.ni : kCoarseMax;
            ca[q][1] = j > 0 ? op.AY(c - op.ni) : 0.f;
            nb[q][2] = i > 0 ? c - 1 : kCoarseMax;
            ca[q][2] = i > 0 ? op.AX(c - 1) : 0.f;
            nb[q][3] = i + 1 < op.ni ? c + 1 : kCoarseMax;
            ca[q][3] = i + 1 < op.ni ? op.AX(c) : 0.f;
            nb[q][4] = j + 1 < op.nj ? c + op.ni : kCoarseMax;
            ca[q][4] = j + 1 < op.nj ? op.AY(c) : 0.f;
            nb[q][5] = k + 1 < op.nk ? c + sl : kCoarseMax;
            ca[q][5] = k + 1 < op.nk ? op.AZ(c) : 0.f;
            xs[c] = 0.f;
        }
    }
    if (threadIdx.x == 0) xs[kCoarseMax] = 0.f;
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
#pragma unroll
        for (int col = 1; col >= 0; --col) {
#pragma unroll
            for (int q = 0; q < kCoarsePer; ++q) {
                if (colour[q] == col) {
                    float s = 0.f;
#pragma unroll
                    for (int e = 0; e < 6; ++e) s += ca[q][e] * xs[nb[q][e]];
                    xs[threadIdx.x + q * kCoarseThreads] = (cd[q] != 0.f) ? (cb[q] - s) / cd[q] : 0.f;
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < kCoarsePer; ++q) {
        int c = threadIdx.x + q * kCoarseThreads;
        if (c < n) x[c] = xs[c];
    }
}

// ----------------------------------------------------------------------------
// Closed-form Galerkin coarsening A_c = 0.5 R A P (R = 1/8, P = injection):
//   coupling(I, I+e) = 1/16 * sum of the fine couplings across the aggregate face
//   leak(I)          = 1/16 * sum of the children's leaks
//   diag(I)          = leak(I) + sum of the 6 coarse couplings' magnitudes
// (equal to 1/16 * sum_{i,k in I} A(i,k); written this way no term cancels).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_coarsen(int ni, int nj, int nk, const float* __restrict__ ax,
                                                    const float* __restrict__ ay, const float* __restrict__ az,
                                                    const float* __restrict__ leak, int ci, int cj, int ck,
                                                    float* __restrict__ dc, float* __restrict__ axc,
                                                    float* __restrict__ ayc, float* __restrict__ azc,
                                                    float* __restrict__ leakc) {
    const size_t nc = (size_t)ci * cj * ck;
    size_t C = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (C >= nc) return;
    int I, J, K;
    unflat(C, ci, cj, I, J, K);
    const float s = 0.0625f;
    // coupling across the aggregate face between fine planes p and p+1 along an axis
    auto face = [&](int axis, int plane) -> float {
        // plane: fine index of the lower side of the face
        float acc = 0.f;
        for (int u = 0; u < 2; ++u)
            for (int v = 0; v < 2; ++v) {
                int i, j, k;
                if (axis == 0) { i = plane; j = 2 * J + u; k = 2 * K + v; }
                else if (axis == 1) { i = 2 * I + u; j = plane; k = 2 * K + v; }
                else { i = 2 * I + u; j = 2 * J + v; k = plane; }
                if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) continue;
                size_t c = lin3(i, j, k, ni, nj);
                acc += (axis == 0 ? ax[c] : axis == 1 ? ay[c] : az[c]);
            }
        return s * acc;
    };
    float cxp = (I + 1 < ci) ? face(0, 2 * I + 1) : 0.f;
    float cyp = (J + 1 < cj) ? face(1, 2 * J + 1) : 0.f;
    float czp = (K + 1 < ck) ? face(2, 2 * K + 1) : 0.f;
    float cxm = (I > 0) ? face(0, 2 * I - 1) : 0.f;
    float cym = (J > 0) ? face(1, 2 * J - 1) : 0.f;
    float czm = (K > 0) ? face(2, 2 * K - 1) : 0.f;
    float lk = 0.f;
    for (int kk = 0; kk < 2; ++kk)
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
                int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                if (i < ni && j < nj && k < nk) lk += leak[lin3(i, j, k, ni, nj)];
            }
    lk *= s;
    leakc[C] = lk;
    axc[C] = cxp;
    ayc[C] = cyp;
    azc[C] = czp;
    dc[C] = lk + ((-czm) + (-cym) + (-cxm) + (-cxp) + (-cyp) + (-czp));
}

// leak = diag - sum|couplings| evaluated in FP64 (for operators given without it).
__global__ __launch_bounds__(kBlock) void k_leak_from_diag(int ni, int nj, int nk, const float* __restrict__ d,
                                                           const float* __restrict__ ax, const float* __restrict__ ay,
                                                           const float* __restrict__ az, float* __restrict__ leak) {
    const size_t n = (size_t)ni * nj * nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat(c, ni, nj, i, j, k);
    const size_t sl = (size_t)ni * nj;
    double s = 0;
    if (k > 0) s -= az[c - sl];
    if (j > 0) s -= ay[c - ni];
    if (i > 0) s -= ax[c - 1];
    if (i + 1 < ni) s -= ax[c];
    if (j + 1 < nj) s -= ay[c];
    if (k + 1 < nk) s -= az[c];
    double l = (double)d[c] - s;
    leak[c] = (float)(l > 0 ? l : 0.0);
}

// FP16 copy of a level scaled by 1/scale.
__global__ __launch_bounds__(kBlock) void k_to_half(size_t n, const float* __restrict__ src, __half* __restrict__ dst,
                                                    float inv_scale) {
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) dst[c] = __float2half(src[c] * inv_scale);
}

// ----------------------------------------------------------------------------
// Krylov kernels. Reductions: per-block partials (grid-stride, fixed grid), then
// one single-block finalize kernel computes the scalars on the device. Acc is
// float (plain FP32 mode) or double (mixed mode: FP64 dots and alpha/beta).
// ----------------------------------------------------------------------------
constexpr int kRedBlocks = 1024;

struct Scalars {
    double rho, rho_new, sigma, alpha, beta, res, bnorm, tol_abs, true_res;
    int breakdown;
};

template <typename CT, typename Acc>
__global__ __launch_bounds__(kBlock) void k_matvec_dot(Op<CT> op, const float* __restrict__ s, float* __restrict__ q,
                                                       double* __restrict__ partial) {
    const size_t n = (size_t)op.ni * op.nj * op.nk;
    Acc acc = 0;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock) {
        int i, j, k;
        unflat(c, op.ni, op.nj, i, j, k);
        const size_t sl = (size_t)op.ni * op.nj;
        // reference multiply order: -z, -y, -x, diag, +x, +y, +z
        float y = 0.f;
        if (k > 0) y += op.AZ(c - sl) * s[c - sl];
        if (j > 0) y += op.AY(c - op.ni) * s[c - op.ni];
        if (i > 0) y += op.AX(c - 1) * s[c - 1];
        y += op.D(c) * s[c];
        if (i + 1 < op.ni) y += op.AX(c) * s[c + 1];
        if (j + 1 < op.nj) y += op.AY(c) * s[c + op.ni];
        if (k + 1 < op.nk) y += op.AZ(c) * s[c + sl];
        q[c] = y;
        acc += (Acc)s[c] * (Acc)y;
    }
    Acc t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_dot(size_t n, const float* __restrict__ a, const float* __restrict__ b,
                                                double* __restrict__ partial) {
    Acc acc = 0;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock)
        acc += (Acc)a[c] * (Acc)b[c];
    Acc t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

__global__ __launch_bounds__(kBlock) void k_absmax(size_t n, const float* __restrict__ a, double* __restrict__ partial) {
    float m = 0.f;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock)
        m = fmaxf(m, fabsf(a[c]));
    float t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

// x += alpha s; r -= alpha q; partial max|r|
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_xr(size_t n, const Scalars* __restrict__ sc, float* __restrict__ x,
                                                      const float* __restrict__ s, float* __restrict__ r,
                                                      const float* __restrict__ q, double* __restrict__ partial) {
    const Acc alpha = (Acc)sc->alpha;
    float m = 0.f;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock) {
        x[c] = (float)((Acc)x[c] + alpha * (Acc)s[c]);
        float rv = (float)((Acc)r[c] - alpha * (Acc)q[c]);
        r[c] = rv;
        m = fmaxf(m, fabsf(rv));
    }
    float t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

// s = z + beta s
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_s(size_t n, const Scalars* __restrict__ sc, float* __restrict__ s,
                                                     const float* __restrict__ z) {
    const Acc beta = (Acc)sc->beta;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock)
        s[c] = (float)((Acc)z[c] + beta * (Acc)s[c]);
}

// r = b - A x evaluated in FP64 (true residual), partial max|r| / |b|.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_true_residual(Op<CT> op, const float* __restrict__ b,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          double* __restrict__ partial) {
    const size_t n = (size_t)op.ni * op.nj * op.nk;
    double m = 0;
    for (size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock) {
        int i, j, k;
        unflat(c, op.ni, op.nj, i, j, k);
        const size_t sl = (size_t)op.ni * op.nj;
        double y = (double)op.D(c) * x[c];
        if (k > 0) y += (double)op.AZ(c - sl) * x[c - sl];
        if (j > 0) y += (double)op.AY(c - op.ni) * x[c - op.ni];
        if (i > 0) y += (double)op.AX(c - 1) * x[c - 1];
        if (i + 1 < op.ni) y += (double)op.AX(c) * x[c + 1];
        if (j + 1 < op.nj) y += (double)op.AY(c) * x[c + op.ni];
        if (k + 1 < op.nk) y += (double)op.AZ(c) * x[c + sl];
        double rv = (double)b[c] - y;
        if (r) r[c] = (float)rv;
        m = fmax(m, fabs(rv));
    }
    double t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Finalize kernels (single block).
template <typename Acc> __device__ inline double reduce_partials_sum(const double* p, int np) {
    Acc a = 0;
    for (int q = threadIdx.x; q < np; q += kBlock) a += (Acc)p[q];
    return (double)block_sum(a);
}
__device__ inline double reduce_partials_max(const double* p, int np) {
    double a = 0;
    for (int q = threadIdx.x; q < np; q += kBlock) a = fmax(a, p[q]);
    return block_max(a);
}

// what: 0 rho (initial) -> also s/beta unused; 1 sigma -> alpha; 2 res; 3 rho_new -> beta;
// 4 bnorm (with tol factor); 5 true_res
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_finalize(Scalars* sc, const double* __restrict__ partial, int np, int what,
                                                     double tol) {
    double v = (what == 2 || what == 4 || what == 5) ? reduce_partials_max(partial, np)
                                                     : reduce_partials_sum<Acc>(partial, np);
    if (threadIdx.x != 0) return;
    if (what == 0) {
        sc->rho = v;
        sc->breakdown = (v == 0 || v != v) ? 1 : 0;
    } else if (what == 1) {
        sc->sigma = v;
        sc->alpha = (double)((Acc)sc->rho / (Acc)v);
    } else if (what == 2) {
        sc->res = v;
    } else if (what == 3) {
        sc->rho_new = v;
        sc->beta = (double)((Acc)v / (Acc)sc->rho);
        sc->rho = v;
    } else if (what == 4) {
        sc->bnorm = v;
        sc->tol_abs = tol * v;
        sc->res = v;
    } else {
        sc->true_res = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill(size_t n, float* __restrict__ a, float v) {
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) a[c] = v;
}
__global__ __launch_bounds__(kBlock) void k_copy(size_t n, float* __restrict__ dst, const float* __restrict__ src) {
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < n) dst[c] = src[c];
}

}  // namespace vflip

// ============================================================================
// Host side
// ============================================================================
using namespace vflip;

struct LevelBuf {
    int ni = 0, nj = 0, nk = 0;
    size_t n = 0;
    float *d = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *leak = nullptr;
    __half *hd = nullptr, *hax = nullptr, *hay = nullptr, *haz = nullptr;
    float hscale = 1.f;
    bool half = false;
    float *x = nullptr, *b = nullptr;
};

struct vflip_solver {
    int ni, nj, nk;
    size_t n;
    vflip_solve_opts opts;
    hipStream_t stream = nullptr;
    std::vector<LevelBuf> L;
    float *r = nullptr, *z = nullptr, *s = nullptr, *q = nullptr, *xw = nullptr;
    double* partial = nullptr;
    Scalars* sc = nullptr;
    Scalars* h_sc = nullptr;  // pinned
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool setup_done = false;
};

namespace {

template <typename CT> Op<CT> make_op(const LevelBuf& l);
template <> Op<float> make_op<float>(const LevelBuf& l) {
    return Op<float>{l.d, l.ax, l.ay, l.az, 1.f, l.ni, l.nj, l.nk};
}
template <> Op<__half> make_op<__half>(const LevelBuf& l) {
    return Op<__half>{l.hd, l.hax, l.hay, l.haz, l.hscale, l.ni, l.nj, l.nk};
}

int alloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return VFLIP_ERR_NOMEM;
    }
    return 0;
}

template <typename CT> int smooth_level(vflip_solver* S, int l, bool pre, const float* xc) {
    LevelBuf& lv = S->L[l];
    Op<CT> op = make_op<CT>(lv);
    const float* b = (l == 0) ? S->r : lv.b;
    float* x = (l == 0) ? S->z : lv.x;
    unsigned g = grid_for(lv.n);
    int sweeps = pre ? S->opts.pre_sweeps : S->opts.post_sweeps;
    int start = 0;
    if (pre) {
        if (sweeps == 0) {
            k_fill<<<g, kBlock, 0, S->stream>>>(lv.n, x, 0.f);
            VF_CHECK_LAUNCH();
            return 0;
        }
        k_rbgs_first<CT><<<g, kBlock, 0, S->stream>>>(op, b, x);
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 0);
        start = 1;
    } else {
        const LevelBuf& c = S->L[l + 1];
        if (sweeps == 0) {
            k_prolong_add<<<g, kBlock, 0, S->stream>>>(lv.ni, lv.nj, lv.nk, x, xc, c.ni, c.nj);
            VF_CHECK_LAUNCH();
            return 0;
        }
        k_rbgs_odd_prolong<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, xc, c.ni, c.nj);
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 0);
        start = 1;
    }
    for (int it = start; it < sweeps; ++it) {
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 1);
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 0);
    }
    VF_CHECK_LAUNCH();
    return 0;
}

template <typename CT> int restrict_level(vflip_solver* S, int l) {
    LevelBuf& lv = S->L[l];
    LevelBuf& c = S->L[l + 1];
    const float* b = (l == 0) ? S->r : lv.b;
    const float* x = (l == 0) ? S->z : lv.x;
    k_residual_restrict<CT><<<grid_for(c.n), kBlock, 0, S->stream>>>(make_op<CT>(lv), b, x, c.b, c.ni, c.nj, c.nk);
    VF_CHECK_LAUNCH();
    return 0;
}

int run_vcycle(vflip_solver* S) {
    const int nl = (int)S->L.size();
    int rc;
    for (int l = 0; l < nl - 1; ++l) {
        rc = S->L[l].half ? smooth_level<__half>(S, l, true, nullptr) : smooth_level<float>(S, l, true, nullptr);
        if (rc) return rc;
        rc = S->L[l].half ? restrict_level<__half>(S, l) : restrict_level<float>(S, l);
        if (rc) return rc;
    }
    {
        LevelBuf& lv = S->L[nl - 1];
        const float* b = (nl == 1) ? S->r : lv.b;
        float* x = (nl == 1) ? S->z : lv.x;
        if (lv.half)
            k_coarsest<__half><<<1, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), b, x, S->opts.coarse_sweeps);
        else
            k_coarsest<float><<<1, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), b, x, S->opts.coarse_sweeps);
        VF_CHECK_LAUNCH();
    }
    for (int l = nl - 2; l >= 0; --l) {
        const float* xc = S->L[l + 1].x;
        rc = S->L[l].half ? smooth_level<__half>(S, l, false, xc) : smooth_level<float>(S, l, false, xc);
        if (rc) return rc;
    }
    return 0;
}

bool mixed(const vflip_solver* S) { return S->opts.precision == VFLIP_PREC_MIXED; }

int finalize(vflip_solver* S, int what, double tol = 0) {
    if (mixed(S))
        k_finalize<double><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial, kRedBlocks, what, tol);
    else
        k_finalize<float><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial, kRedBlocks, what, tol);
    VF_CHECK_LAUNCH();
    return 0;
}

int fetch_scalars(vflip_solver* S) {
    VF_HIP(hipMemcpyAsync(S->h_sc, S->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int check_dims(int ni, int nj, int nk) {
    if (ni < 1 || nj < 1 || nk < 1) return 0;
    if ((double)ni * nj * nk > 2.0e9) return 0;
    return 1;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void vflip_default_opts(vflip_solve_opts* o) {
    if (!o) return;
    o->tolerance = 1e-6;
    o->max_iterations = 50;
    o->precision = VFLIP_PREC_FP32;
    o->pre_sweeps = 4;
    o->post_sweeps = 4;
    o->coarse_sweeps = 200;
    o->refine_every = 8;
    o->warm_start = 0;
}

int vflip_solver_destroy(vflip_solver* S) {
    if (!S) return 0;
    for (auto& l : S->L) {
        if (l.d) (void)hipFree(l.d);
        if (l.hd) (void)hipFree(l.hd);
        if (l.x) (void)hipFree(l.x);
    }
    if (S->r) (void)hipFree(S->r);
    if (S->partial) (void)hipFree(S->partial);
    if (S->sc) (void)hipFree(S->sc);
    if (S->h_sc) (void)hipHostFree(S->h_sc);
    if (S->e0) (void)hipEventDestroy(S->e0);
    if (S->e1) (void)hipEventDestroy(S->e1);
    delete S;
    return 0;
}

int vflip_solver_set_opts(vflip_solver* S, const vflip_solve_opts* o) {
    VF_ARG(S && o, "solver/opts");
    VF_ARG(o->max_iterations >= 0 && o->pre_sweeps >= 0 && o->post_sweeps >= 0 && o->coarse_sweeps >= 0,
           "negative iteration or sweep count");
    VF_ARG(o->precision == VFLIP_PREC_FP32 || o->precision == VFLIP_PREC_MIXED, "precision");
    bool need_half = (o->precision == VFLIP_PREC_MIXED);
    bool had_half = (S->opts.precision == VFLIP_PREC_MIXED);
    S->opts = *o;
    if (S->opts.refine_every < 0) S->opts.refine_every = 0;
    if (need_half != had_half) {
        for (size_t l = 0; l < S->L.size(); ++l) S->L[l].half = need_half && l >= 2 && S->L[l].hd;
        S->setup_done = false;
    }
    return 0;
}

int vflip_solver_create(int ni, int nj, int nk, const vflip_solve_opts* opts, vflip_solver** out) {
    VF_ARG(out, "out");
    *out = nullptr;
    VF_ARG(check_dims(ni, nj, nk), "grid dims");
    vflip_solver* S = new vflip_solver();
    S->ni = ni; S->nj = nj; S->nk = nk;
    S->n = (size_t)ni * nj * nk;
    vflip_default_opts(&S->opts);
    if (opts) S->opts = *opts;
    // level shapes, S/GeometricLevelGen.h:36-58
    {
        int a = ni, b = nj, c = nk;
        unsigned unknowns = (unsigned)ni * nj * nk;
        S->L.push_back(LevelBuf());
        S->L.back().ni = a; S->L.back().nj = b; S->L.back().nk = c;
        while (unknowns > 16u * 16u * 16u) {
            a = (int)std::ceil((float)a / 2.0);
            b = (int)std::ceil((float)b / 2.0);
            c = (int)std::ceil((float)c / 2.0);
            S->L.push_back(LevelBuf());
            S->L.back().ni = a; S->L.back().nj = b; S->L.back().nk = c;
            unknowns = (unsigned)a * b * c;
        }
    }
    int rc = 0;
    for (size_t l = 0; l < S->L.size() && !rc; ++l) {
        LevelBuf& lv = S->L[l];
        lv.n = (size_t)lv.ni * lv.nj * lv.nk;
        void* p = nullptr;
        rc = alloc(&p, 5 * lv.n * sizeof(float));
        if (rc) break;
        lv.d = (float*)p; lv.ax = lv.d + lv.n; lv.ay = lv.ax + lv.n; lv.az = lv.ay + lv.n; lv.leak = lv.az + lv.n;
        if (l >= 1) {
            rc = alloc(&p, 2 * lv.n * sizeof(float));
            if (rc) break;
            lv.x = (float*)p; lv.b = lv.x + lv.n;
        }
        if (l >= 2) {
            rc = alloc(&p, 4 * lv.n * sizeof(__half));
            if (rc) break;
            lv.hd = (__half*)p; lv.hax = lv.hd + lv.n; lv.hay = lv.hax + lv.n; lv.haz = lv.hay + lv.n;
            lv.half = (S->opts.precision == VFLIP_PREC_MIXED);
        }
    }
    if (!rc) {
        void* p = nullptr;
        rc = alloc(&p, 5 * S->n * sizeof(float));
        if (!rc) {
            S->r = (float*)p; S->z = S->r + S->n; S->s = S->z + S->n; S->q = S->s + S->n; S->xw = S->q + S->n;
        }
    }
    if (!rc) rc = alloc((void**)&S->partial, kRedBlocks * sizeof(double));
    if (!rc) rc = alloc((void**)&S->sc, sizeof(Scalars));
    if (!rc) {
        if (hipHostMalloc((void**)&S->h_sc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
            set_error("hipHostMalloc failed");
            rc = VFLIP_ERR_NOMEM;
        }
    }
    if (!rc) {
        if (hipEventCreate(&S->e0) != hipSuccess || hipEventCreate(&S->e1) != hipSuccess) {
            set_error("hipEventCreate failed");
            rc = VFLIP_ERR_HIP;
        }
    }
    if (rc) {
        vflip_solver_destroy(S);
        return rc;
    }
    *out = S;
    return 0;
}

int vflip_solver_set_stream(vflip_solver* S, void* stream) {
    VF_ARG(S, "solver");
    S->stream = (hipStream_t)stream;
    return 0;
}

int vflip_solver_levels(const vflip_solver* S) { return S ? (int)S->L.size() : 0; }

int vflip_solver_level_dims(const vflip_solver* S, int l, int* ni, int* nj, int* nk) {
    VF_ARG(S && l >= 0 && l < (int)S->L.size(), "level");
    if (ni) *ni = S->L[l].ni;
    if (nj) *nj = S->L[l].nj;
    if (nk) *nk = S->L[l].nk;
    return 0;
}

int vflip_solver_set_matrix(vflip_solver* S, const float* d, const float* ax, const float* ay, const float* az,
                            const float* leak) {
    VF_ARG(S && d && ax && ay && az, "matrix arrays");
    LevelBuf& l0 = S->L[0];
    size_t bytes = l0.n * sizeof(float);
    VF_HIP(hipMemcpyAsync(l0.d, d, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.ax, ax, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.ay, ay, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.az, az, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (leak) {
        VF_HIP(hipMemcpyAsync(l0.leak, leak, bytes, hipMemcpyDeviceToDevice, S->stream));
    } else {
        k_leak_from_diag<<<grid_for(l0.n), kBlock, 0, S->stream>>>(l0.ni, l0.nj, l0.nk, l0.d, l0.ax, l0.ay, l0.az,
                                                                   l0.leak);
        VF_CHECK_LAUNCH();
    }
    S->setup_done = false;
    return 0;
}

int vflip_solver_setup(vflip_solver* S) {
    VF_ARG(S, "solver");
    VF_HIP(hipEventRecord(S->e0, S->stream));
    for (size_t l = 0; l + 1 < S->L.size(); ++l) {
        LevelBuf& f = S->L[l];
        LevelBuf& c = S->L[l + 1];
        k_coarsen<<<grid_for(c.n), kBlock, 0, S->stream>>>(f.ni, f.nj, f.nk, f.ax, f.ay, f.az, f.leak, c.ni, c.nj,
                                                          c.nk, c.d, c.ax, c.ay, c.az, c.leak);
        VF_CHECK_LAUNCH();
    }
    if (S->opts.precision == VFLIP_PREC_MIXED) {
        // FP16 copies of levels >= 2, scaled so the largest diagonal maps to 2^10.
        for (size_t l = 2; l < S->L.size(); ++l) {
            LevelBuf& lv = S->L[l];
            k_absmax<<<kRedBlocks, kBlock, 0, S->stream>>>(lv.n, lv.d, S->partial);
            VF_CHECK_LAUNCH();
            int rc = finalize(S, 5);
            if (rc) return rc;
            rc = fetch_scalars(S);
            if (rc) return rc;
            double mx = S->h_sc->true_res;
            float scale = (mx > 0) ? (float)(mx / 1024.0) : 1.f;
            lv.hscale = scale;
            float inv = 1.f / scale;
            k_to_half<<<grid_for(lv.n), kBlock, 0, S->stream>>>(lv.n, lv.d, lv.hd, inv);
            k_to_half<<<grid_for(lv.n), kBlock, 0, S->stream>>>(lv.n, lv.ax, lv.hax, inv);
            k_to_half<<<grid_for(lv.n), kBlock, 0, S->stream>>>(lv.n, lv.ay, lv.hay, inv);
            k_to_half<<<grid_for(lv.n), kBlock, 0, S->stream>>>(lv.n, lv.az, lv.haz, inv);
            VF_CHECK_LAUNCH();
            lv.half = true;
        }
    } else {
        for (auto& lv : S->L) lv.half = false;
    }
    VF_HIP(hipEventRecord(S->e1, S->stream));
    S->setup_done = true;
    return 0;
}

int vflip_solver_get_level(const vflip_solver* S, int l, float* d, float* ax, float* ay, float* az) {
    VF_ARG(S && l >= 0 && l < (int)S->L.size(), "level");
    const LevelBuf& lv = S->L[l];
    size_t bytes = lv.n * sizeof(float);
    if (d) VF_HIP(hipMemcpyAsync(d, lv.d, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (ax) VF_HIP(hipMemcpyAsync(ax, lv.ax, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (ay) VF_HIP(hipMemcpyAsync(ay, lv.ay, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (az) VF_HIP(hipMemcpyAsync(az, lv.az, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_vcycle(vflip_solver* S, const float* r, float* z) {
    VF_ARG(S && r && z, "solver/r/z");
    if (!S->setup_done) {
        int rc = vflip_solver_setup(S);
        if (rc) return rc;
    }
    VF_HIP(hipMemcpyAsync(S->r, r, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    int rc = run_vcycle(S);
    if (rc) return rc;
    VF_HIP(hipMemcpyAsync(z, S->z, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_matvec(vflip_solver* S, const float* x, float* y) {
    VF_ARG(S && x && y, "solver/x/y");
    k_matvec_dot<float, float><<<kRedBlocks, kBlock, 0, S->stream>>>(make_op<float>(S->L[0]), x, y, S->partial);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_solve(vflip_solver* S, const float* rhs, float* x, vflip_solve_stats* st) {
    VF_ARG(S && rhs && x, "solver/rhs/x");
    int rc;
    float ms_setup = 0.f;
    if (!S->setup_done) {
        rc = vflip_solver_setup(S);
        if (rc) return rc;
        VF_HIP(hipEventSynchronize(S->e1));
        VF_HIP(hipEventElapsedTime(&ms_setup, S->e0, S->e1));
    }
    const size_t n = S->n;
    const bool mx = mixed(S);
    Op<float> A = make_op<float>(S->L[0]);
    VF_HIP(hipEventRecord(S->e0, S->stream));
    // |b|_inf and the stop threshold (S/AlgebraicMultigrid.h:212-233)
    k_absmax<<<kRedBlocks, kBlock, 0, S->stream>>>(n, rhs, S->partial);
    VF_CHECK_LAUNCH();
    if ((rc = finalize(S, 4, S->opts.tolerance))) return rc;
    if (S->opts.warm_start) {
        VF_HIP(hipMemcpyAsync(S->xw, x, n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
        k_true_residual<float><<<kRedBlocks, kBlock, 0, S->stream>>>(A, rhs, S->xw, S->r, S->partial);
        VF_CHECK_LAUNCH();
        if ((rc = finalize(S, 2))) return rc;
    } else {
        k_fill<<<grid_for(n), kBlock, 0, S->stream>>>(n, S->xw, 0.f);
        k_copy<<<grid_for(n), kBlock, 0, S->stream>>>(n, S->r, rhs);
        VF_CHECK_LAUNCH();
    }
    if ((rc = fetch_scalars(S))) return rc;
    vflip_solve_stats stats{};
    stats.ms_setup = ms_setup;
    double bnorm = S->h_sc->bnorm;
    double tol_abs = S->h_sc->tol_abs;
    double res = S->h_sc->res;
    int status = VFLIP_OK;
    int it = 0;
    if (bnorm == 0) {
        k_fill<<<grid_for(n), kBlock, 0, S->stream>>>(n, x, 0.f);
        VF_CHECK_LAUNCH();
        stats.converged = 1;
        res = 0;
    } else if (res <= tol_abs && S->opts.warm_start) {
        stats.converged = 1;
    } else {
        if ((rc = run_vcycle(S))) return rc;
        if (mx) k_dot<double><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->z, S->r, S->partial);
        else k_dot<float><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->z, S->r, S->partial);
        VF_CHECK_LAUNCH();
        if ((rc = finalize(S, 0))) return rc;
        if ((rc = fetch_scalars(S))) return rc;
        if (S->h_sc->breakdown) {
            status = VFLIP_ERR_BREAKDOWN;
        } else {
            k_copy<<<grid_for(n), kBlock, 0, S->stream>>>(n, S->s, S->z);
            VF_CHECK_LAUNCH();
            status = VFLIP_NOT_CONVERGED;
            for (it = 0; it < S->opts.max_iterations; ++it) {
                if (mx) k_matvec_dot<float, double><<<kRedBlocks, kBlock, 0, S->stream>>>(A, S->s, S->q, S->partial);
                else k_matvec_dot<float, float><<<kRedBlocks, kBlock, 0, S->stream>>>(A, S->s, S->q, S->partial);
                VF_CHECK_LAUNCH();
                if ((rc = finalize(S, 1))) return rc;
                if (mx) k_update_xr<double><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->sc, S->xw, S->s, S->r, S->q, S->partial);
                else k_update_xr<float><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->sc, S->xw, S->s, S->r, S->q, S->partial);
                VF_CHECK_LAUNCH();
                if ((rc = finalize(S, 2))) return rc;
                if (mx && S->opts.refine_every > 0 && ((it + 1) % S->opts.refine_every) == 0) {
                    k_true_residual<float><<<kRedBlocks, kBlock, 0, S->stream>>>(A, rhs, S->xw, S->r, S->partial);
                    VF_CHECK_LAUNCH();
                    if ((rc = finalize(S, 2))) return rc;
                }
                if ((rc = fetch_scalars(S))) return rc;
                res = S->h_sc->res;
                if (res <= tol_abs) {
                    status = VFLIP_OK;
                    ++it;
                    break;
                }
                if ((rc = run_vcycle(S))) return rc;
                if (mx) k_dot<double><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->z, S->r, S->partial);
                else k_dot<float><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->z, S->r, S->partial);
                VF_CHECK_LAUNCH();
                if ((rc = finalize(S, 3))) return rc;
                if (mx) k_update_s<double><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->sc, S->s, S->z);
                else k_update_s<float><<<kRedBlocks, kBlock, 0, S->stream>>>(n, S->sc, S->s, S->z);
                VF_CHECK_LAUNCH();
            }
            stats.converged = (status == VFLIP_OK);
        }
        VF_HIP(hipMemcpyAsync(x, S->xw, n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    }
    VF_HIP(hipEventRecord(S->e1, S->stream));
    // FP64 true residual of the returned solution, for the stats
    k_true_residual<float><<<kRedBlocks, kBlock, 0, S->stream>>>(A, rhs, x, nullptr, S->partial);
    VF_CHECK_LAUNCH();
    if ((rc = finalize(S, 5))) return rc;
    if ((rc = fetch_scalars(S))) return rc;
    float ms = 0.f;
    VF_HIP(hipEventElapsedTime(&ms, S->e0, S->e1));
    stats.ms_solve = ms;
    stats.iterations = it;
    stats.residual = res;
    stats.rel_residual = bnorm > 0 ? res / bnorm : 0.0;
    stats.true_residual = bnorm > 0 ? S->h_sc->true_res / bnorm : 0.0;
    if (st) *st = stats;
    return status;
}

int vflip_amgpcg_solve_host(int ni, int nj, int nk, const double* diag, const double* ax, const double* ay,
                            const double* az, const double* rhs, double* result, double tol, int maxit,
                            double* residual_out, int* iterations_out, const vflip_solve_opts* opts) {
    VF_ARG(diag && ax && ay && az && rhs && result, "host arrays");
    VF_ARG(check_dims(ni, nj, nk), "grid dims");
    vflip_solve_opts o;
    vflip_default_opts(&o);
    if (opts) o = *opts;
    o.tolerance = tol;
    o.max_iterations = maxit;
    o.warm_start = 0;
    vflip_solver* S = nullptr;
    int rc = vflip_solver_create(ni, nj, nk, &o, &S);
    if (rc) return rc;
    const size_t n = (size_t)ni * nj * nk;
    std::vector<float> h(6 * n);
    float *hd = h.data(), *hx = hd + n, *hy = hx + n, *hz = hy + n, *hl = hz + n, *hb = hl + n;
    for (size_t c = 0; c < n; ++c) {
        hd[c] = (float)diag[c];
        hx[c] = (float)ax[c];
        hy[c] = (float)ay[c];
        hz[c] = (float)az[c];
        hb[c] = (float)rhs[c];
    }
    // leak in FP64: diag - sum of |couplings| of the row
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                size_t c = lin3(i, j, k, ni, nj);
                double s = 0;
                if (i > 0) s -= ax[c - 1];
                if (j > 0) s -= ay[c - ni];
                if (k > 0) s -= az[c - (size_t)ni * nj];
                s -= ax[c] + ay[c] + az[c];
                double l = diag[c] - s;
                hl[c] = (float)(l > 0 ? l : 0.0);
            }
    float* dev = nullptr;
    if (hipMalloc((void**)&dev, 7 * n * sizeof(float)) != hipSuccess) {
        vflip_solver_destroy(S);
        set_error("hipMalloc failed");
        return VFLIP_ERR_NOMEM;
    }
    vflip_solve_stats st{};
    int status = 0;
    do {
        if (hipMemcpy(dev, h.data(), 6 * n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("H2D copy failed");
            status = VFLIP_ERR_HIP;
            break;
        }
        status = vflip_solver_set_matrix(S, dev, dev + n, dev + 2 * n, dev + 3 * n, dev + 4 * n);
        if (status < 0) break;
        status = vflip_solver_solve(S, dev + 5 * n, dev + 6 * n, &st);
        if (status < 0 && status != VFLIP_ERR_BREAKDOWN) break;
        std::vector<float> xo(n);
        if (hipMemcpy(xo.data(), dev + 6 * n, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("D2H copy failed");
            status = VFLIP_ERR_HIP;
            break;
        }
        for (size_t c = 0; c < n; ++c) result[c] = xo[c];
        if (residual_out) *residual_out = st.residual;
        if (iterations_out) *iterations_out = st.iterations;
    } while (0);
    (void)hipFree(dev);
    vflip_solver_destroy(S);
    return status;
}

}  // extern "C"

namespace vflip {
int solver_level0(vflip_solver* S, int* ni, int* nj, int* nk, float** d, float** ax, float** ay, float** az,
                  float** leak, hipStream_t* stream) {
    VF_ARG(S, "solver");
    LevelBuf& l = S->L[0];
    *ni = l.ni; *nj = l.nj; *nk = l.nk;
    *d = l.d; *ax = l.ax; *ay = l.ay; *az = l.az; *leak = l.leak;
    *stream = S->stream;
    return 0;
}
void solver_mark_matrix_changed(vflip_solver* S) { S->setup_done = false; }
}  // namespace vflip
